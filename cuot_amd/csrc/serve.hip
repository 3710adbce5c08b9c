// serve.hip -- the serving loop: correlation requests of different sizes are
// admitted against an HBM budget and run with continuous batching.
//
// One scheduler thread owns the device arena (one allocation of the budget,
// carved by a first-fit allocator on the host) and S streams.  Whenever a
// stream is free and requests are queued, it takes queued requests in arrival
// order while their outputs plus the batch workspace fit in the free arena,
// plans them as one cuot_batch and launches it; a host function enqueued
// behind the batch marks its requests complete, so batches on different
// streams overlap on the device and new arrivals are admitted as soon as a
// stream and memory free up.  A request's outputs stay in the arena until the
// client releases them; batch workspaces are returned on completion.
//
// plain = 1 turns the same loop into the comparator: one stream, one request
// at a time, each run by cuot_sot_run / cuot_ferret_run (FIFO plain loop).
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <list>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "ctx.h"
#include "engine.h"
#include "kernels.h"

namespace {

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

size_t a256(size_t v) { return (v + 255) & ~(size_t)255; }

// first-fit allocator over [0, size) with coalescing free list
struct Arena {
  size_t size = 0;
  std::map<size_t, size_t> free_;  // offset -> length
  size_t used = 0;
  void init(size_t s) {
    size = s;
    free_.clear();
    free_[0] = s;
    used = 0;
  }
  bool alloc(size_t len, size_t* off) {
    len = a256(len);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= len) {
        *off = it->first;
        const size_t rest = it->second - len;
        const size_t noff = it->first + len;
        free_.erase(it);
        if (rest) free_[noff] = rest;
        used += len;
        return true;
      }
    }
    return false;
  }
  void release(size_t off, size_t len) {
    len = a256(len);
    used -= len;
    auto it = free_.emplace(off, len).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
};

struct Req {
  cuot_instance inst;
  double t_submit = 0, t_start = 0, t_done = 0;
  size_t out_off = 0, out_bytes = 0;
  int state = 0;  // 0 queued, 1 running, 2 done
  int status = CUOT_OK;
};

size_t out_bytes_of(const cuot_instance& in) {
  const size_t n = (size_t)in.n;
  return a256(n * 16) * 2 + a256(((n + 31) / 32) * 4) + 256;
}

struct Flight;

// COT relation of one served result: SOT w_i = v_i ^ b_i * Delta (b packed
// LSB-first), Ferret r_i = s_i ^ lsb(r_i) * Delta with lsb(Delta) = 1.  Each
// mismatching index adds one to *fail.
__global__ void relation_check_kernel(const uint4* __restrict__ o0, const uint4* __restrict__ o1,
                                      const uint32_t* __restrict__ bits, const uint4* __restrict__ delta, int64_t n,
                                      int kind, unsigned long long* fail) {
  const uint4 d = delta[0];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint4 a = o0[i], b = o1[i];
    const bool sel = kind == CUOT_MODE_SOT ? ((bits[i >> 5] >> (i & 31)) & 1u) : (b.x & 1u);
    const uint4 e = sel ? make_uint4(a.x ^ d.x, a.y ^ d.y, a.z ^ d.z, a.w ^ d.w) : a;
    const bool bad = e.x != b.x || e.y != b.y || e.z != b.z || e.w != b.w || (kind != CUOT_MODE_SOT && i == 0 && !(d.x & 1u));
    if (bad) atomicAdd(fail, 1ull);
  }
}

}  // namespace

struct cuot_server {
  cuot_ctx* ctx = nullptr;
  cuot_server_config cfg;
  char* arena_d = nullptr;
  bool own_arena = true;  // false: the caller owns arena_d (cuot_server_config.arena_d)
  bool pool_streams = false;  // the streams are the context's worker streams (not destroyed here)
  Arena arena;
  std::vector<hipStream_t> streams;
  std::vector<bool> busy;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<int64_t> queue;
  std::unordered_map<int64_t, Req> reqs;
  int64_t next_ticket = 1;
  std::list<Flight*> inflight;
  std::thread worker;
  bool stop = false;
  int64_t launched_batches = 0;
  // resident Ferret clients (cuot_server_config.ferret_clients): an iterating
  // session each, plus the outputs of its last extend (sender n_ext blocks,
  // then receiver n_ext blocks) of which `left` usable COTs from `pos` on have
  // not been handed out yet.  pos/left are touched by the scheduler thread
  // only; busy is guarded by mu.
  struct Client {
    cuot_ferret_session* sess = nullptr;
    char* buf = nullptr;
    int64_t n_ext = 0, ot_limit = 0, pos = 0, left = 0;
    bool busy = false;
  };
  std::vector<Client> clients;

  char* out0(const Req& r) const { return arena_d + r.out_off; }
  char* out1(const Req& r) const { return arena_d + r.out_off + a256((size_t)r.inst.n * 16); }
  char* bits(const Req& r) const { return arena_d + r.out_off + 2 * a256((size_t)r.inst.n * 16); }
  char* delta(const Req& r) const {
    return arena_d + r.out_off + 2 * a256((size_t)r.inst.n * 16) + a256((((size_t)r.inst.n + 31) / 32) * 4);
  }
  void loop();
  int launch_batch(std::unique_lock<std::mutex>& lk, int s);
  int launch_client(std::unique_lock<std::mutex>& lk, int s, int64_t ticket);
  void erase_queued(int64_t ticket) {
    auto it = std::find(queue.begin(), queue.end(), ticket);
    if (it != queue.end()) queue.erase(it);
  }
};

namespace {

struct Flight {
  cuot_server* srv = nullptr;
  cuot_batch* batch = nullptr;
  std::vector<int64_t> tickets;
  size_t ws_off = 0, ws_bytes = 0;
  int stream = 0;
  int client = 0;  // > 0: a resident-client request (the client is busy until the flight is reaped)
  std::atomic<bool> done{false};
  double t_done = 0;
  int status = CUOT_OK;
};

void on_done(void* p) {
  Flight* f = (Flight*)p;
  f->t_done = now_s();
  {
    std::lock_guard<std::mutex> g(f->srv->mu);
    f->done.store(true);
  }
  f->srv->cv.notify_all();
}

// plan + workspace size of the instance list (host only)
int plan_bytes(cuot_ctx* ctx, const std::vector<cuot_instance>& v, size_t chunk, cuot_batch** b, size_t* ws) {
  int rc = cuot_batch_new(ctx, v.data(), (int)v.size(), chunk, b);
  if (rc) return rc;
  *ws = cuot_batch_workspace_bytes(*b);
  return CUOT_OK;
}

int plain_ws_bytes(const cuot_instance& in, size_t* ws) {
  if (in.kind == CUOT_MODE_SOT) {
    cuot_sot_config c;
    memset(&c, 0, sizeof(c));
    c.n = in.n;
    c.t = in.t;
    c.h = in.h;
    c.k0 = in.k0;
    c.k1 = in.k1;
    c.seed = in.seed;
    *ws = cuot_sot_workspace_bytes(&c);
  } else {
    cuot_ferret_config c;
    memset(&c, 0, sizeof(c));
    c.n = in.n;
    c.k = in.k;
    c.t = in.t;
    c.h = in.h;
    c.ntasks = in.ntasks < 1 ? 1 : in.ntasks;
    c.seed = in.seed;
    *ws = cuot_ferret_workspace_bytes(&c);
  }
  if (*ws == 0) {
    cuot_set_last_error("invalid instance", __FILE__, __LINE__);
    return CUOT_ERR_ARG;
  }
  return CUOT_OK;
}

}  // namespace

// Called with the lock held; takes requests from the queue head.  Returns
// 1 if a batch was launched, 0 if nothing could be admitted.
int cuot_server::launch_batch(std::unique_lock<std::mutex>& lk, int s) {
  // The first queued request that can start decides: a resident-client
  // request whose client is idle runs alone on the client's session; a
  // request without client starts a batch of the queued requests without
  // client (in arrival order).  Requests of a busy client wait their turn.
  for (int64_t t : queue) {
    Req& r = reqs[t];
    if (r.inst.client == 0) break;
    if (r.inst.client < 0 || r.inst.client > (int)clients.size() || r.inst.kind != CUOT_MODE_FERRET) {
      r.state = 2;
      r.status = CUOT_ERR_ARG;
      r.t_done = now_s();
      erase_queued(t);
      cv.notify_all();
      return 1;
    }
    if (!clients[r.inst.client - 1].busy) return launch_client(lk, s, t);
  }
  std::vector<int64_t> take;
  std::vector<cuot_instance> insts;
  cuot_batch* b = nullptr;
  size_t ws = 0, outs = 0;
  const int maxb = cfg.plain ? 1 : (cfg.max_batch > 0 ? cfg.max_batch : 64);
  const int64_t max_cots = cfg.max_batch_cots > 0 ? cfg.max_batch_cots : ((int64_t)1 << 24);
  int64_t cots = 0;
  for (int64_t t : queue) {
    if ((int)take.size() >= maxb) break;
    Req& r = reqs[t];
    if (r.inst.client != 0) continue;  // served by its client's session
    if (!take.empty() && cots + r.inst.n > max_cots) break;  // keep batches short
    std::vector<cuot_instance> trial = insts;
    trial.push_back(r.inst);
    size_t tws = 0;
    cuot_batch* tb = nullptr;
    int rc = cfg.plain ? plain_ws_bytes(r.inst, &tws) : plan_bytes(ctx, trial, cfg.qc_chunk_bytes, &tb, &tws);
    if (rc) {
      if (take.empty()) {  // a request that cannot be planned fails alone
        r.state = 2;
        r.status = rc;
        r.t_done = now_s();
        erase_queued(t);
        cv.notify_all();
        return 1;
      }
      break;
    }
    const size_t tout = outs + out_bytes_of(r.inst);
    // fit test against the free arena: outputs of every request + workspace
    Arena probe = arena;
    bool fits = true;
    size_t off;
    for (size_t j = 0; j <= take.size() && fits; j++) {
      const cuot_instance& in = j < take.size() ? reqs[take[j]].inst : r.inst;
      fits = probe.alloc(out_bytes_of(in), &off);
    }
    if (fits) fits = probe.alloc(tws, &off);
    if (!fits) {
      if (tb) cuot_batch_free(tb);
      if (take.empty() && inflight.empty() && tout + tws > arena.size) {
        r.state = 2;
        r.status = CUOT_ERR_NOMEM;
        r.t_done = now_s();
        erase_queued(t);
        cv.notify_all();
        return 1;
      }
      break;
    }
    if (b) cuot_batch_free(b);
    b = tb;
    ws = tws;
    outs = tout;
    cots += r.inst.n;
    take.push_back(t);
    insts.swap(trial);
  }
  if (take.empty()) return 0;
  Flight* f = new Flight();
  f->srv = this;
  f->batch = b;
  f->tickets = take;
  f->stream = s;
  f->ws_bytes = ws;
  for (int64_t t : take) {
    Req& r = reqs[t];
    arena.alloc(out_bytes_of(r.inst), &r.out_off);
    r.out_bytes = out_bytes_of(r.inst);
    r.state = 1;
    r.t_start = now_s();
  }
  for (int64_t t : take) erase_queued(t);
  arena.alloc(ws, &f->ws_off);
  busy[s] = true;
  inflight.push_back(f);
  launched_batches++;
  // launch outside the lock (host-side launch work overlaps new submits)
  std::vector<cuot_block*> o0, o1, dl;
  std::vector<uint8_t*> bt;
  for (int64_t t : take) {
    const Req& r = reqs[t];
    o0.push_back((cuot_block*)out0(r));
    o1.push_back((cuot_block*)out1(r));
    bt.push_back((uint8_t*)bits(r));
    dl.push_back((cuot_block*)delta(r));
  }
  const cuot_instance first = reqs[take[0]].inst;
  lk.unlock();
  hipStream_t st = streams[s];
  char* wsp = arena_d + f->ws_off;
  int rc = CUOT_OK;
  if (cfg.plain) {
    if (first.kind == CUOT_MODE_SOT) {
      cuot_sot_config c;
      memset(&c, 0, sizeof(c));
      c.n = first.n;
      c.t = first.t;
      c.h = first.h;
      c.k0 = first.k0;
      c.k1 = first.k1;
      c.seed = first.seed;
      rc = cuot_sot_run(ctx, &c, wsp, ws, o0[0], o1[0], bt[0], dl[0], st);
    } else {
      cuot_ferret_config c;
      memset(&c, 0, sizeof(c));
      c.n = first.n;
      c.k = first.k;
      c.t = first.t;
      c.h = first.h;
      c.ntasks = first.ntasks < 1 ? 1 : first.ntasks;
      c.seed = first.seed;
      rc = cuot_ferret_run(ctx, &c, wsp, ws, o0[0], o1[0], dl[0], st);
    }
  } else {
    rc = cuot_batch_bind_on(b, wsp, ws, o0.data(), o1.data(), bt.data(), dl.data(), st);
    if (!rc) rc = cuot_batch_run(b, cfg.use_graph ? (CUOT_BATCH_STAGE_ALL | CUOT_BATCH_GRAPH) : CUOT_BATCH_STAGE_ALL,
                                 st);
  }
  if (!rc && hipLaunchHostFunc(st, on_done, f) != hipSuccess) rc = CUOT_ERR_HIP;
  lk.lock();
  if (rc) {
    // nothing was enqueued behind the work that may have been launched:
    // wait for the stream, then fail the requests
    lk.unlock();
    hipStreamSynchronize(st);
    lk.lock();
    f->status = rc;
    f->t_done = now_s();
    f->done.store(true);
  }
  return 1;
}

// A resident-client request (called with the lock held): its outputs are
// taken from the arena, then on stream s the client's unused COTs are copied
// and further extends run until n COTs are handed out (FerretCOT::rcot,
// ferret_cot.hpp:171-209).  Returns 1 if launched (or failed), 0 if the arena
// has no room yet.
int cuot_server::launch_client(std::unique_lock<std::mutex>& lk, int s, int64_t t) {
  Req& r = reqs[t];
  Client& cl = clients[r.inst.client - 1];
  const size_t ob = out_bytes_of(r.inst);
  size_t off = 0;
  if (!arena.alloc(ob, &off)) {
    if (inflight.empty() && ob > arena.size) {
      r.state = 2;
      r.status = CUOT_ERR_NOMEM;
      r.t_done = now_s();
      erase_queued(t);
      cv.notify_all();
      return 1;
    }
    return 0;
  }
  r.out_off = off;
  r.out_bytes = ob;
  r.state = 1;
  r.t_start = now_s();
  erase_queued(t);
  Flight* f = new Flight();
  f->srv = this;
  f->tickets.push_back(t);
  f->stream = s;
  f->client = r.inst.client;
  cl.busy = true;
  busy[s] = true;
  inflight.push_back(f);
  launched_batches++;
  const int64_t n = r.inst.n;
  char* o0 = out0(r);
  char* o1 = out1(r);
  char* dl = delta(r);
  lk.unlock();
  hipStream_t st = streams[s];
  int rc = CUOT_OK;
  int64_t done = 0;
  while (!rc && done < n) {
    if (cl.left == 0) {
      rc = cuot_ferret_session_extend(cl.sess, (cuot_block*)cl.buf, (cuot_block*)(cl.buf + cl.n_ext * 16), st);
      cl.pos = 0;
      cl.left = rc ? 0 : cl.ot_limit;
      if (rc) break;
    }
    const int64_t take = std::min(cl.left, n - done);
    if (hipMemcpyAsync(o0 + done * 16, cl.buf + cl.pos * 16, (size_t)take * 16, hipMemcpyDeviceToDevice, st) ||
        hipMemcpyAsync(o1 + done * 16, cl.buf + (cl.n_ext + cl.pos) * 16, (size_t)take * 16,
                       hipMemcpyDeviceToDevice, st))
      rc = CUOT_ERR_HIP;
    cl.pos += take;
    cl.left -= take;
    done += take;
  }
  if (!rc) rc = cuot_ferret_session_delta(cl.sess, (cuot_block*)dl, st);
  if (!rc && hipLaunchHostFunc(st, on_done, f) != hipSuccess) rc = CUOT_ERR_HIP;
  lk.lock();
  if (rc) {
    lk.unlock();
    hipStreamSynchronize(st);
    lk.lock();
    f->status = rc;
    f->t_done = now_s();
    f->done.store(true);
  }
  return 1;
}

void cuot_server::loop() {
  hipSetDevice(ctx->device);
  std::unique_lock<std::mutex> lk(mu);
  for (;;) {
    // reap finished batches; their plans (graphs, device tables) are freed
    // after the lock is dropped, so that no HIP call runs while on_done may be
    // waiting for the lock
    std::vector<Flight*> finished;
    for (auto it = inflight.begin(); it != inflight.end();) {
      Flight* f = *it;
      if (!f->done.load()) {
        ++it;
        continue;
      }
      for (int64_t t : f->tickets) {
        auto rit = reqs.find(t);
        if (rit == reqs.end()) continue;
        rit->second.state = 2;
        rit->second.status = f->status;
        rit->second.t_done = f->t_done;
      }
      if (f->ws_bytes) arena.release(f->ws_off, f->ws_bytes);
      if (f->client > 0) clients[f->client - 1].busy = false;
      busy[f->stream] = false;
      finished.push_back(f);
      it = inflight.erase(it);
    }
    if (!finished.empty()) {
      cv.notify_all();
      lk.unlock();
      for (Flight* f : finished) {
        if (f->batch) cuot_batch_free(f->batch);
        delete f;
      }
      lk.lock();
    }
    if (stop && inflight.empty()) {
      // requests still queued at shutdown fail, and their waiters wake
      const double td = now_s();
      for (int64_t t : queue) {
        auto rit = reqs.find(t);
        if (rit == reqs.end()) continue;
        rit->second.state = 2;
        rit->second.status = CUOT_ERR_STATE;
        rit->second.t_done = td;
      }
      queue.clear();
      cv.notify_all();
      break;
    }
    bool launched = false;
    if (!stop) {
      for (size_t s = 0; s < streams.size() && !queue.empty(); s++) {
        if (busy[s]) continue;
        if (launch_batch(lk, (int)s)) launched = true;
        break;  // re-scan state after each launch
      }
    }
    if (launched) continue;
    cv.wait_for(lk, std::chrono::milliseconds(2));
  }
}

extern "C" int cuot_server_new(cuot_ctx* ctx, const cuot_server_config* cfg, cuot_server** out) {
  CUOT_CHECK_ARG(ctx && cfg && out, "NULL argument");
  CUOT_CHECK_ARG(cfg->hbm_budget >= (1u << 20), "hbm_budget below 1 MiB");
  CUOT_CHECK_ARG(cfg->streams >= 0 && cfg->streams <= 8, "streams must be in [0, 8]");
  CUOT_CHECK_ARG(cfg->ferret_clients >= 0 && cfg->ferret_clients <= 64, "ferret_clients must be in [0, 64]");
  *out = nullptr;
  cuot_server* s = new cuot_server();
  s->ctx = ctx;
  s->cfg = *cfg;
  const int ns = cfg->plain ? 1 : (cfg->streams > 0 ? cfg->streams : 2);
  if (hipSetDevice(ctx->device) != hipSuccess) {
    cuot_set_last_error("hipSetDevice failed", __FILE__, __LINE__);
    delete s;
    return CUOT_ERR_HIP;
  }
  auto free_clients = [s] {
    for (auto& cl : s->clients) {
      if (cl.sess) cuot_ferret_session_free(cl.sess);
      if (cl.buf) hipFree(cl.buf);
    }
    s->clients.clear();
  };
  // resident Ferret clients: their device memory counts against the budget
  size_t sess_bytes = 0;
  if (cfg->ferret_clients > 0) {
    size_t f0 = 0, f1 = 0, tot = 0;
    hipMemGetInfo(&f0, &tot);
    for (int c = 1; c <= cfg->ferret_clients; c++) {
      cuot_ferret_config fc = cfg->ferret_session;
      fc.seed.lo ^= (uint64_t)c;
      cuot_server::Client cl;
      int rc = cuot_ferret_session_new(ctx, &fc, &cl.sess);
      if (!rc) rc = cuot_ferret_session_info(cl.sess, nullptr, &cl.ot_limit, nullptr);
      cl.n_ext = fc.n;
      if (!rc && hipMalloc((void**)&cl.buf, (size_t)fc.n * 32) != hipSuccess) {
        cuot_set_last_error("cannot allocate a client's extend buffer", __FILE__, __LINE__);
        rc = CUOT_ERR_NOMEM;
      }
      s->clients.push_back(cl);
      if (rc) {
        free_clients();
        delete s;
        return rc;
      }
    }
    hipMemGetInfo(&f1, &tot);
    sess_bytes = f0 > f1 ? f0 - f1 : 0;
    if (sess_bytes + (1u << 20) > cfg->hbm_budget) {
      free_clients();
      delete s;
      cuot_set_last_error("hbm_budget does not hold the resident Ferret clients", __FILE__, __LINE__);
      return CUOT_ERR_NOMEM;
    }
  }
  const size_t arena_bytes = cfg->hbm_budget - sess_bytes;
  s->own_arena = cfg->arena_d == nullptr;
  if (!s->own_arena) {
    s->arena_d = (char*)cfg->arena_d;
  } else if (hipMalloc((void**)&s->arena_d, arena_bytes) != hipSuccess) {
    free_clients();
    cuot_set_last_error("cannot allocate the HBM budget", __FILE__, __LINE__);
    delete s;
    return CUOT_ERR_NOMEM;
  }
  s->arena.init(arena_bytes);
  // the context's worker streams when free (they hold distinct hardware
  // queues); otherwise streams of the server's own
  s->pool_streams = ns <= 2 && ctx->work[0] && !ctx->work_taken.exchange(true);
  for (int i = 0; i < ns; i++) {
    hipStream_t st = s->pool_streams ? ctx->work[i] : nullptr;
    if (!st && hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
      if (!s->pool_streams)
        for (hipStream_t x : s->streams) hipStreamDestroy(x);
      else
        ctx->work_taken = false;
      if (s->own_arena) hipFree(s->arena_d);
      free_clients();
      delete s;
      cuot_set_last_error("stream creation failed", __FILE__, __LINE__);
      return CUOT_ERR_HIP;
    }
    s->streams.push_back(st);
    s->busy.push_back(false);
  }
  CUOT_HIP_TRY(cuot_qc_configure());
  s->worker = std::thread([s] { s->loop(); });
  *out = s;
  return CUOT_OK;
}

extern "C" int cuot_server_submit(cuot_server* s, const cuot_instance* inst, int64_t* ticket) {
  CUOT_CHECK_ARG(s && inst && ticket, "NULL argument");
  CUOT_CHECK_ARG(inst->kind == CUOT_MODE_SOT || inst->kind == CUOT_MODE_FERRET, "bad instance kind");
  CUOT_CHECK_ARG(inst->n >= 1, "n < 1");
  CUOT_CHECK_ARG(inst->client == 0 || (inst->kind == CUOT_MODE_FERRET && inst->client > 0 &&
                                       inst->client <= (int)s->clients.size()),
                 "client must be 0 or a resident Ferret client (Ferret requests only)");
  {
    std::lock_guard<std::mutex> g(s->mu);
    if (s->stop) {
      cuot_set_last_error("server is shutting down", __FILE__, __LINE__);
      return CUOT_ERR_STATE;
    }
    const int64_t t = s->next_ticket++;
    Req r;
    r.inst = *inst;
    r.t_submit = now_s();
    s->reqs.emplace(t, r);
    s->queue.push_back(t);
    *ticket = t;
  }
  s->cv.notify_all();
  return CUOT_OK;
}

extern "C" int cuot_server_poll(cuot_server* s, int64_t ticket, int wait, cuot_server_result* res) {
  CUOT_CHECK_ARG(s && res, "NULL argument");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  CUOT_CHECK_ARG(it != s->reqs.end(), "unknown ticket");
  // a ticket released by another thread meanwhile counts as finished
  if (wait) {
    s->cv.wait(lk, [&] {
      auto w = s->reqs.find(ticket);
      return w == s->reqs.end() || w->second.state == 2;
    });
    it = s->reqs.find(ticket);
    CUOT_CHECK_ARG(it != s->reqs.end(), "ticket released while waiting");
  }
  const Req& r = it->second;
  memset(res, 0, sizeof(*res));
  res->done = r.state == 2;
  res->status = r.status;
  res->t_submit = r.t_submit;
  res->t_start = r.t_start;
  res->t_done = r.t_done;
  if (r.state == 2 && r.status == CUOT_OK) {
    res->out0 = (cuot_block*)s->out0(r);
    res->out1 = (cuot_block*)s->out1(r);
    res->bits = r.inst.kind == CUOT_MODE_SOT ? (uint8_t*)s->bits(r) : nullptr;
    res->delta = (cuot_block*)s->delta(r);
  }
  return CUOT_OK;
}

extern "C" int cuot_server_release(cuot_server* s, int64_t ticket) {
  CUOT_CHECK_ARG(s, "NULL server");
  {
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(ticket);
    CUOT_CHECK_ARG(it != s->reqs.end(), "unknown ticket");
    if (it->second.state != 2) {
      cuot_set_last_error("request still queued or running", __FILE__, __LINE__);
      return CUOT_ERR_BUSY;
    }
    if (it->second.out_bytes) s->arena.release(it->second.out_off, it->second.out_bytes);
    s->reqs.erase(it);
  }
  s->cv.notify_all();
  return CUOT_OK;
}

extern "C" int cuot_server_stats(cuot_server* s, int64_t* batches, size_t* arena_used) {
  CUOT_CHECK_ARG(s, "NULL server");
  std::lock_guard<std::mutex> g(s->mu);
  if (batches) *batches = s->launched_batches;
  if (arena_used) *arena_used = s->arena.used;
  return CUOT_OK;
}

extern "C" int cuot_server_free(cuot_server* s) {
  if (!s) return CUOT_OK;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
  }
  s->cv.notify_all();
  if (s->worker.joinable()) s->worker.join();
  if (s->pool_streams)
    s->ctx->work_taken = false;
  else
    for (hipStream_t st : s->streams) hipStreamDestroy(st);
  if (s->own_arena) hipFree(s->arena_d);
  for (auto& cl : s->clients) {
    cuot_ferret_session_free(cl.sess);
    hipFree(cl.buf);
  }
  delete s;
  return CUOT_OK;
}

// Replays an arrival schedule in real time (request i is submitted at
// arrival_s[i] seconds after the start), waits for every request, releases
// its outputs and reports its latency (completion - arrival) in latency_s.
// With relation_failures != NULL every completed result is checked on the
// device (relation_check_kernel, on a stream of its own, after its completion
// time was taken) before its outputs are released; *relation_failures is the
// number of indices, over all results, that break the COT relation.
extern "C" int cuot_serve_replay(cuot_ctx* ctx, const cuot_server_config* cfg, const cuot_instance* reqs,
                                 const double* arrival_s, int count, double* latency_s, double* wall_s,
                                 int64_t* batches, int64_t* relation_failures) {
  CUOT_CHECK_ARG(ctx && cfg && reqs && arrival_s && latency_s && count >= 1, "bad argument");
  cuot_server* s = nullptr;
  int rc = cuot_server_new(ctx, cfg, &s);
  if (rc) return rc;
  const bool check = relation_failures != nullptr;
  hipStream_t cst = nullptr;
  unsigned long long* fail_d = nullptr;
  const bool own_cst = !s->pool_streams;  // the context's third worker stream checks pooled servers
  if (check) {
    if (!own_cst) cst = ctx->work[2];
    if ((own_cst && hipStreamCreateWithFlags(&cst, hipStreamNonBlocking) != hipSuccess) ||
        hipMalloc((void**)&fail_d, sizeof(unsigned long long)) != hipSuccess ||
        hipMemsetAsync(fail_d, 0, sizeof(unsigned long long), cst) != hipSuccess) {
      if (fail_d) hipFree(fail_d);
      if (cst && own_cst) hipStreamDestroy(cst);
      cuot_server_free(s);
      cuot_set_last_error("relation check set-up failed", __FILE__, __LINE__);
      return CUOT_ERR_HIP;
    }
  }
  std::vector<int64_t> tickets(count, 0);
  std::vector<double> arrive(count, 0);
  std::vector<int> pending;
  std::vector<std::pair<int64_t, hipEvent_t>> checking;  // released once their check has run
  const double t0 = now_s();
  int first_err = CUOT_OK;
  // record latency and status of a finished request, then check and release it
  auto finish = [&](int j, const cuot_server_result& r) {
    latency_s[j] = r.t_done - arrive[j];
    if (r.status && !first_err) first_err = r.status;
    if (check && r.status == CUOT_OK) {
      const int64_t n = reqs[j].n;
      const int blocks = (int)std::min<int64_t>(2048, (n + 255) / 256);
      hipLaunchKernelGGL(relation_check_kernel, dim3(blocks), dim3(256), 0, cst, (const uint4*)r.out0,
                         (const uint4*)r.out1, (const uint32_t*)r.bits, (const uint4*)r.delta, n, reqs[j].kind,
                         fail_d);
      hipEvent_t ev = nullptr;
      if (hipGetLastError() != hipSuccess || hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess ||
          hipEventRecord(ev, cst) != hipSuccess) {
        if (!first_err) first_err = CUOT_ERR_HIP;
        hipStreamSynchronize(cst);
        if (ev) hipEventDestroy(ev);
        cuot_server_release(s, tickets[j]);
        return;
      }
      checking.emplace_back(tickets[j], ev);
    } else {
      cuot_server_release(s, tickets[j]);
    }
  };
  auto release_checked = [&](bool all) {
    for (size_t q = 0; q < checking.size();) {
      if (all ? hipEventSynchronize(checking[q].second) == hipSuccess
              : hipEventQuery(checking[q].second) == hipSuccess) {
        hipEventDestroy(checking[q].second);
        cuot_server_release(s, checking[q].first);
        checking[q] = checking.back();
        checking.pop_back();
      } else {
        q++;
      }
    }
  };
  for (int i = 0; i < count; i++) {
    const double due = t0 + arrival_s[i];
    double tn = now_s();
    if (due > tn) std::this_thread::sleep_for(std::chrono::duration<double>(due - tn));
    arrive[i] = t0 + arrival_s[i];
    rc = cuot_server_submit(s, &reqs[i], &tickets[i]);
    if (rc) {
      first_err = rc;
      count = i;
      break;
    }
    pending.push_back(i);
    // reap what has finished meanwhile so the arena can be reused
    for (size_t q = 0; q < pending.size();) {
      const int j = pending[q];
      cuot_server_result r;
      if (cuot_server_poll(s, tickets[j], 0, &r) == CUOT_OK && r.done) {
        finish(j, r);
        tickets[j] = -tickets[j];
        pending[q] = pending.back();
        pending.pop_back();
      } else {
        q++;
      }
    }
    release_checked(false);
  }
  for (int j = 0; j < count; j++) {
    if (tickets[j] <= 0) continue;
    cuot_server_result r;
    cuot_server_poll(s, tickets[j], 1, &r);
    finish(j, r);
    release_checked(false);
  }
  if (wall_s) *wall_s = now_s() - t0;
  release_checked(true);
  if (check) {
    unsigned long long fails = 0;
    if (hipMemcpy(&fails, fail_d, sizeof(fails), hipMemcpyDeviceToHost) != hipSuccess && !first_err)
      first_err = CUOT_ERR_HIP;
    *relation_failures = (int64_t)fails;
    hipFree(fail_d);
    if (own_cst) hipStreamDestroy(cst);
  }
  if (batches) cuot_server_stats(s, batches, nullptr);
  cuot_server_free(s);
  return first_err;
}
