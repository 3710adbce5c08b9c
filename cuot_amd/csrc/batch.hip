// batch.hip -- many SOT and Ferret instances as one launch sequence.
//
// Every instance keeps its own workspace carve (the SotPlan / FerretPlan of a
// single run), so its stage outputs are the same bytes as a plain run.  What
// is shared is the launches:
//   * GGM: one launch per (role, subtree split L) over the trees of all
//     instances (instance tables + tree prefix in HBM);
//   * A^ prepare: one launch sequence per transform length 2^m;
//   * QC encode: one launch sequence per 2^m and scratch chunk, both parties of
//     every instance of the class in it (gridDim.y = party);
//   * LPN: one launch over the rows of all Ferret instances.
// Per-instance work that is only a few blocks long (seed PRG streams, masked
// base-OT rows, level-sum bookkeeping, sparse encode) keeps its own launches;
// with CUOT_BATCH_GRAPH the whole sequence is a captured HIP graph.
#include <string.h>

#include <algorithm>
#include <map>
#include <vector>

#include "ctx.h"
#include "engine.h"
#include "kernels.h"

namespace {

constexpr size_t kDefaultQcChunk = (size_t)4 << 30;

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

struct GgmGroup {
  int L = 0;
  bool recv = false;
  std::vector<int> members;
  int64_t total = 0;  // trees
  size_t tab_off = 0, start_off = 0;
};

struct PrepGroup {
  int m = 0;
  std::vector<int> members;
  size_t off = 0;
};

struct QcChunk {
  int m = 0;
  std::vector<std::pair<int, int>> parties;  // (instance, 0 sender / 1 receiver)
  int64_t nmax = 0;                          // largest n of the chunk
  size_t io_off = 0, ahat_off = 0;
  bool aux = false;  // second stream under CUOT_BATCH_OVERLAP (own scratch region)
};

}  // namespace

struct cuot_batch {
  cuot_ctx* ctx = nullptr;
  std::vector<cuot_instance> inst;
  std::vector<SotPlan> sot;
  std::vector<FerretPlan> fer;
  std::vector<size_t> inst_off;
  std::vector<GgmGroup> ggm;
  std::vector<PrepGroup> prep;
  std::vector<QcChunk> qc;
  std::vector<int> ferret_ids;
  size_t lpn_tab_off = 0, lpn_start_off = 0;
  int64_t lpn_wgs = 0;
  size_t tables_bytes = 0, qc_scratch_off = 0, qc_scratch_bytes = 0, ws_bytes = 0;
  size_t qc_scratch2_off = 0;  // scratch of the chunks marked aux
  size_t qc_chunk_bytes = kDefaultQcChunk;
  // bound state
  char* ws = nullptr;
  std::vector<blk*> out0, out1;
  std::vector<uint32_t*> bits;
  std::vector<blk*> delta;  // per-instance Delta destination (may be NULL)
  std::vector<char> img;    // host image of the launch tables
  bool bound = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint32_t graph_stages = 0;

  bool is_sot(int i) const { return inst[i].kind == CUOT_MODE_SOT; }
  void drop_graph() {
    if (exec) hipGraphExecDestroy(exec);
    if (graph) hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    graph_stages = 0;
  }
};

namespace {

int plan_batch(cuot_batch* b) {
  const int count = (int)b->inst.size();
  b->sot.assign(count, SotPlan());
  b->fer.assign(count, FerretPlan());
  // ---- per-instance plans and validation --------------------------------
  for (int i = 0; i < count; i++) {
    const cuot_instance& in = b->inst[i];
    if (in.kind == CUOT_MODE_SOT) {
      SotPlan& p = b->sot[i];
      memset(&p.cfg, 0, sizeof(p.cfg));
      p.cfg.n = in.n;
      p.cfg.t = in.t;
      p.cfg.h = in.h;
      p.cfg.k0 = in.k0;
      p.cfg.k1 = in.k1;
      p.cfg.seed = in.seed;
      p.shared_qc = true;
      int rc = p.validate();
      if (rc) return rc;
    } else if (in.kind == CUOT_MODE_FERRET) {
      FerretPlan& p = b->fer[i];
      memset(&p.cfg, 0, sizeof(p.cfg));
      p.cfg.n = in.n;
      p.cfg.k = in.k;
      p.cfg.t = in.t;
      p.cfg.h = in.h;
      p.cfg.ntasks = in.ntasks < 1 ? 1 : in.ntasks;
      p.cfg.seed = in.seed;
      int rc = p.validate();
      if (rc) return rc;
      b->ferret_ids.push_back(i);
    } else {
      cuot_set_last_error("instance kind must be CUOT_MODE_SOT or CUOT_MODE_FERRET", __FILE__, __LINE__);
      return CUOT_ERR_ARG;
    }
  }
  // ---- groups ------------------------------------------------------------
  std::map<int, int> ggm_s, ggm_r;  // L -> group index
  std::map<int, std::vector<int>> by_m;
  for (int i = 0; i < count; i++) {
    const int L = cuot_ggm_split(b->inst[i].h);
    for (int r = 0; r < 2; r++) {
      std::map<int, int>& mp = r ? ggm_r : ggm_s;
      auto it = mp.find(L);
      if (it == mp.end()) {
        GgmGroup g;
        g.L = L;
        g.recv = r == 1;
        b->ggm.push_back(g);
        it = mp.emplace(L, (int)b->ggm.size() - 1).first;
      }
      GgmGroup& g = b->ggm[it->second];
      g.members.push_back(i);
      g.total += b->inst[i].t;
    }
    if (b->is_sot(i)) by_m[cuot_qc_log_len(b->inst[i].n)].push_back(i);
  }
  // senders before receivers
  std::stable_sort(b->ggm.begin(), b->ggm.end(), [](const GgmGroup& x, const GgmGroup& y) {
    return x.recv != y.recv ? !x.recv : x.L < y.L;
  });
  for (auto& kv : by_m) {
    PrepGroup pg;
    pg.m = kv.first;
    pg.members = kv.second;
    b->prep.push_back(pg);
    // chunks of both parties under the scratch cap (one party per chunk when
    // the transform length has no batched form)
    const int m = kv.first;
    const bool multi = cuot_qc_batchable(m);
    QcChunk cur;
    cur.m = m;
    for (int i : kv.second) {
      for (int party = 0; party < 2; party++) {
        const int B = (int)cur.parties.size() + 1;
        const bool fits = multi && cuot_qc_batch_scratch_bytes(m, B) <= b->qc_chunk_bytes;
        if (!cur.parties.empty() && !fits) {
          b->qc.push_back(cur);
          cur.parties.clear();
          cur.nmax = 0;
        }
        cur.parties.emplace_back(i, party);
        cur.nmax = std::max(cur.nmax, b->inst[i].n);
      }
    }
    if (!cur.parties.empty()) b->qc.push_back(cur);
  }
  // ---- workspace layout ----------------------------------------------------
  size_t off = 0;
  for (GgmGroup& g : b->ggm) {
    g.tab_off = off;
    off = align256(off + g.members.size() * sizeof(GgmKArgs));
    g.start_off = off;
    off = align256(off + (g.members.size() + 1) * sizeof(int64_t));
    if (g.total > 0x7fffffff) {
      cuot_set_last_error("too many trees in one batch", __FILE__, __LINE__);
      return CUOT_ERR_ARG;
    }
  }
  for (PrepGroup& pg : b->prep) {
    pg.off = off;
    off = align256(off + pg.members.size() * sizeof(QcPrep));
  }
  for (QcChunk& c : b->qc) {
    c.io_off = off;
    off = align256(off + c.parties.size() * sizeof(QcIo));
    c.ahat_off = off;
    off = align256(off + c.parties.size() * sizeof(void*));
  }
  if (!b->ferret_ids.empty()) {
    b->lpn_tab_off = off;
    off = align256(off + b->ferret_ids.size() * cuot_lpn_args_bytes());
    b->lpn_start_off = off;
    off = align256(off + (b->ferret_ids.size() + 1) * sizeof(int64_t));
  }
  b->tables_bytes = off;
  b->inst_off.assign(count, 0);
  for (int i = 0; i < count; i++) {
    b->inst_off[i] = off;
    off = align256(off + (b->is_sot(i) ? b->sot[i].carve(nullptr) : b->fer[i].carve(nullptr)));
  }
  // Chunks by decreasing work (parties x M log M) go to the two streams in
  // turn, the largest to the main stream: with CUOT_BATCH_OVERLAP the aux
  // chunks run on the second stream with their own scratch region.
  std::vector<int> order(b->qc.size());
  for (size_t j = 0; j < order.size(); j++) order[j] = (int)j;
  auto work = [&](int j) { return (double)b->qc[j].parties.size() * (double)((int64_t)1 << b->qc[j].m) * b->qc[j].m; };
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return work(x) > work(y); });
  size_t scr1 = 0, scr2 = 0;
  for (size_t r = 0; r < order.size(); r++) {
    QcChunk& c = b->qc[order[r]];
    c.aux = (r & 1) != 0;
    const size_t need = cuot_qc_batch_scratch_bytes(c.m, (int)c.parties.size());
    if (c.aux)
      scr2 = std::max(scr2, need);
    else
      scr1 = std::max(scr1, need);
  }
  b->qc_scratch_bytes = scr1;
  b->qc_scratch_off = off;
  b->qc_scratch2_off = align256(off + scr1);
  b->ws_bytes = b->qc_scratch2_off + scr2 + 256;
  return CUOT_OK;
}

// host image of the launch tables, uploaded once per bind
int build_tables(cuot_batch* b, std::vector<char>& img) {
  img.assign(b->tables_bytes, 0);
  for (const GgmGroup& g : b->ggm) {
    GgmKArgs* tab = (GgmKArgs*)(img.data() + g.tab_off);
    int64_t* start = (int64_t*)(img.data() + g.start_off);
    int64_t acc = 0;
    for (size_t j = 0; j < g.members.size(); j++) {
      const int i = g.members[j];
      GgmKArgs a;
      if (b->is_sot(i)) {
        const SotPlan& p = b->sot[i];
        p.ggm_args(a);
        if (!g.recv) {
          a.roots = p.roots;
          a.leaves = p.leaves_s;
          a.sums = p.sums;
        } else {
          a.leaves = p.leaves_r;
          a.choice = p.choice;
          a.csum = p.csum;
          a.fin = p.fin;
        }
      } else {
        const FerretPlan& p = b->fer[i];
        p.ggm_args(a);
        if (!g.recv) {
          a.roots = p.roots;
          a.leaves = b->out0[i];
          a.sums = p.sums;
        } else {
          a.leaves = b->out1[i];
          a.choice = p.choice;
          a.csum = p.csum;
          a.fin = p.ssum;
        }
      }
      memcpy(&tab[j], &a, sizeof(a));
      start[j] = acc;
      acc += b->inst[i].t;
    }
    start[g.members.size()] = acc;
  }
  for (const PrepGroup& pg : b->prep) {
    QcPrep* tab = (QcPrep*)(img.data() + pg.off);
    for (size_t j = 0; j < pg.members.size(); j++) {
      const SotPlan& p = b->sot[pg.members[j]];
      tab[j].a_bits = (const uint8_t*)p.a_words;
      tab[j].ahat = p.ahat;
      tab[j].n = p.cfg.n;
    }
  }
  for (const QcChunk& c : b->qc) {
    QcIo* io = (QcIo*)(img.data() + c.io_off);
    const uint32_t** ah = (const uint32_t**)(img.data() + c.ahat_off);
    for (size_t j = 0; j < c.parties.size(); j++) {
      const int i = c.parties[j].first;
      const SotPlan& p = b->sot[i];
      const bool recv = c.parties[j].second == 1;
      io[j].x = recv ? p.leaves_r : p.leaves_s;
      io[j].y = recv ? b->out1[i] : b->out0[i];
      io[j].n = p.cfg.n;
      ah[j] = p.ahat;
    }
  }
  if (!b->ferret_ids.empty()) {
    const size_t rec = cuot_lpn_args_bytes();
    int64_t* start = (int64_t*)(img.data() + b->lpn_start_off);
    int64_t acc = 0;
    for (size_t j = 0; j < b->ferret_ids.size(); j++) {
      const int i = b->ferret_ids[j];
      LpnLaunch l;
      b->fer[i].lpn_launch(b->out0[i], b->out1[i], l);
      const int64_t wgs = cuot_lpn_args_fill(l, img.data() + b->lpn_tab_off + j * rec);
      if (wgs < 0) {
        cuot_set_last_error("LPN arguments out of range", __FILE__, __LINE__);
        return CUOT_ERR_ARG;
      }
      start[j] = acc;
      acc += wgs;
    }
    start[b->ferret_ids.size()] = acc;
    b->lpn_wgs = acc;
    if (acc > 0x7fffffff) {
      cuot_set_last_error("too many LPN rows in one batch", __FILE__, __LINE__);
      return CUOT_ERR_ARG;
    }
  }
  return CUOT_OK;
}

// ax (non-null, all stages): the context's auxiliary stream takes the A^
// preparation, the sparse encodes, the Ferret LPN and the QC chunks marked
// aux, forked after the setup and joined at the end (graph capture follows
// the events), so that the LPN gathers and small encodes run beside the large
// encodes.
hipError_t launch_all(cuot_batch* b, uint32_t stages, hipStream_t st, hipStream_t ax) {
  hipError_t e;
  const int count = (int)b->inst.size();
  char* ws = b->ws;
  const bool two = ax != nullptr && stages == CUOT_BATCH_STAGE_ALL;
  hipEvent_t* ev = b->ctx->ev;
  if (stages & CUOT_BATCH_STAGE_SETUP) {
    for (int i = 0; i < count; i++) {
      if ((e = b->is_sot(i) ? b->sot[i].run_setup(st) : b->fer[i].run_setup(st))) return e;
      if (b->delta[i]) {
        const blk* d = b->is_sot(i) ? b->sot[i].delta : b->fer[i].delta;
        if ((e = hipMemcpyAsync(b->delta[i], d, 16, hipMemcpyDeviceToDevice, st))) return e;
      }
    }
  }
  hipStream_t sp = st;  // preparation, sparse encodes, LPN
  if (two) {
    if ((e = hipEventRecord(ev[0], st))) return e;
    if ((e = hipStreamWaitEvent(ax, ev[0], 0))) return e;
    sp = ax;
  }
  if (stages & CUOT_BATCH_STAGE_QC_PREPARE)
    for (const PrepGroup& pg : b->prep)
      if ((e = cuot_launch_qc_prepare_batch((const QcPrep*)(ws + pg.off), (int)pg.members.size(), pg.m, sp)))
        return e;
  if (two) {
    if ((e = hipEventRecord(ev[1], sp))) return e;
  }
  if (stages & CUOT_BATCH_STAGE_SPARSE)
    for (int i = 0; i < count; i++) {
      if (!b->is_sot(i)) continue;
      const SotPlan& p = b->sot[i];
      if ((e = cuot_launch_sparse(p.a_words, p.cfg.n, p.pos, (int)p.cfg.t, b->bits[i], sp))) return e;
    }
  for (int r = 0; r < 2; r++) {
    if (!(stages & (r ? CUOT_BATCH_STAGE_GGM_RECVER : CUOT_BATCH_STAGE_GGM_SENDER))) continue;
    for (const GgmGroup& g : b->ggm) {
      if (g.recv != (r == 1)) continue;
      if ((e = cuot_launch_ggm_batch((const GgmKArgs*)(ws + g.tab_off), (const int64_t*)(ws + g.start_off),
                                     (int)g.members.size(), g.total, g.L, g.recv, st)))
        return e;
    }
    if (r == 0) {
      for (int i = 0; i < count; i++) {
        if (b->is_sot(i)) {
          if ((e = b->sot[i].run_sender_msgs(st))) return e;
        } else {
          if ((e = b->fer[i].run_sender_sums(st))) return e;
        }
      }
    }
  }
  if (two) {
    if ((e = hipEventRecord(ev[2], st))) return e;
    if ((e = hipStreamWaitEvent(ax, ev[2], 0))) return e;
    if ((e = hipStreamWaitEvent(st, ev[1], 0))) return e;
  }
  if ((stages & CUOT_BATCH_STAGE_LPN) && !b->ferret_ids.empty()) {
    for (int i : b->ferret_ids)
      if ((e = b->fer[i].fill_kkp(sp))) return e;
    if ((e = cuot_launch_lpn_batch(ws + b->lpn_tab_off, (const int64_t*)(ws + b->lpn_start_off),
                                   (int)b->ferret_ids.size(), b->lpn_wgs, 2, sp)))
      return e;
  }
  // QC chunks: on one stream all share a scratch region at a time; under the
  // overlap the aux chunks use the second region on the second stream
  if (stages & CUOT_BATCH_STAGE_QC)
    for (const QcChunk& c : b->qc) {
      const bool on_ax = two && c.aux;
      char* scr = ws + (c.aux ? b->qc_scratch2_off : b->qc_scratch_off);
      if ((e = cuot_launch_qc_encode_batch((const QcIo*)(ws + c.io_off), (const uint32_t* const*)(ws + c.ahat_off),
                                           (int)c.parties.size(), c.m, c.nmax, scr, on_ax ? ax : st)))
        return e;
    }
  if (two) {
    if ((e = hipEventRecord(ev[3], ax))) return e;
    if ((e = hipStreamWaitEvent(st, ev[3], 0))) return e;
  }
  return hipSuccess;
}

}  // namespace

extern "C" int cuot_batch_new(cuot_ctx* ctx, const cuot_instance* inst, int count, size_t qc_chunk_bytes,
                              cuot_batch** out) {
  CUOT_CHECK_ARG(ctx && inst && out, "NULL argument");
  CUOT_CHECK_ARG(count >= 1 && count <= (1 << 20), "count out of range");
  *out = nullptr;
  cuot_batch* b = new cuot_batch();
  b->ctx = ctx;
  b->inst.assign(inst, inst + count);
  if (qc_chunk_bytes) b->qc_chunk_bytes = qc_chunk_bytes;
  int rc = plan_batch(b);
  if (rc) {
    delete b;
    return rc;
  }
  *out = b;
  return CUOT_OK;
}

extern "C" size_t cuot_batch_workspace_bytes(const cuot_batch* b) { return b ? b->ws_bytes : 0; }

int cuot_batch_bind_on(cuot_batch* b, void* ws_d, size_t ws_bytes, cuot_block* const* out0, cuot_block* const* out1,
                       uint8_t* const* bits, cuot_block* const* delta, hipStream_t st) {
  CUOT_CHECK_ARG(b && ws_d && out0 && out1, "NULL argument");
  CUOT_CHECK_ARG(ws_bytes >= b->ws_bytes, "workspace too small (cuot_batch_workspace_bytes)");
  CUOT_CHECK_ARG(((uintptr_t)ws_d & 255) == 0, "workspace must be 256-byte aligned");
  const int count = (int)b->inst.size();
  for (int i = 0; i < count; i++) {
    CUOT_CHECK_ARG(out0[i] && out1[i], "output pointer is NULL");
    CUOT_CHECK_ARG(!b->is_sot(i) || (bits && bits[i]), "SOT instance needs a choice-bit buffer");
  }
  b->drop_graph();
  b->bound = false;
  b->ws = (char*)ws_d;
  b->out0.assign(count, nullptr);
  b->out1.assign(count, nullptr);
  b->bits.assign(count, nullptr);
  b->delta.assign(count, nullptr);
  for (int i = 0; i < count; i++) {
    b->out0[i] = (blk*)out0[i];
    b->out1[i] = (blk*)out1[i];
    if (bits) b->bits[i] = (uint32_t*)bits[i];
    if (delta) b->delta[i] = (blk*)delta[i];
    if (b->is_sot(i))
      b->sot[i].carve(b->ws + b->inst_off[i]);
    else
      b->fer[i].carve(b->ws + b->inst_off[i]);
  }
  int rc = build_tables(b, b->img);
  if (rc) return rc;
  CUOT_HIP_TRY(hipSetDevice(b->ctx->device));
  CUOT_HIP_TRY(cuot_qc_configure());
  if (!b->img.empty()) {
    // the image stays alive in the batch until it is freed, so an async copy
    // from it is safe
    if (st)
      CUOT_HIP_TRY(hipMemcpyAsync(b->ws, b->img.data(), b->img.size(), hipMemcpyHostToDevice, st));
    else
      CUOT_HIP_TRY(hipMemcpy(b->ws, b->img.data(), b->img.size(), hipMemcpyHostToDevice));
  }
  b->bound = true;
  return CUOT_OK;
}

extern "C" int cuot_batch_bind(cuot_batch* b, void* ws_d, size_t ws_bytes, cuot_block* const* out0,
                               cuot_block* const* out1, uint8_t* const* bits, cuot_block* delta_d) {
  CUOT_CHECK_ARG(b, "NULL batch");
  std::vector<cuot_block*> d(b->inst.size(), nullptr);
  if (delta_d)
    for (size_t i = 0; i < d.size(); i++) d[i] = delta_d + i;
  return cuot_batch_bind_on(b, ws_d, ws_bytes, out0, out1, bits, d.data(), nullptr);
}

extern "C" int cuot_batch_run(cuot_batch* b, uint32_t stages, void* stream) {
  CUOT_CHECK_ARG(b, "NULL batch");
  if (!b->bound) {
    cuot_set_last_error("batch is not bound (cuot_batch_bind)", __FILE__, __LINE__);
    return CUOT_ERR_STATE;
  }
  CUOT_HIP_TRY(hipSetDevice(b->ctx->device));
  hipStream_t st = b->ctx->st(stream);
  const uint32_t work = stages & CUOT_BATCH_STAGE_ALL;
  hipStream_t ax = nullptr;
  if (stages & CUOT_BATCH_OVERLAP) {
    CUOT_HIP_TRY(b->ctx->ensure_aux());
    ax = b->ctx->aux;
  }
  if (!(stages & CUOT_BATCH_GRAPH)) {
    CUOT_HIP_TRY(launch_all(b, work, st, ax));
    return CUOT_OK;
  }
  CUOT_CHECK_ARG(st != nullptr, "graph replay needs a stream (not the null stream)");
  const uint32_t key = work | (stages & CUOT_BATCH_OVERLAP);
  if (!b->exec || b->graph_stages != key) {
    b->drop_graph();
    CUOT_HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    hipError_t e = launch_all(b, work, st, ax);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(st, &g);
    if (e == hipSuccess) e = e2;
    if (e != hipSuccess) {
      if (g) hipGraphDestroy(g);
      CUOT_HIP_TRY(e);
    }
    b->graph = g;
    hipError_t e3 = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    if (e3 != hipSuccess) {
      b->drop_graph();
      CUOT_HIP_TRY(e3);
    }
    b->graph_stages = key;
  }
  CUOT_HIP_TRY(hipGraphLaunch(b->exec, st));
  return CUOT_OK;
}

extern "C" int cuot_batch_free(cuot_batch* b) {
  if (!b) return CUOT_OK;
  b->drop_graph();
  delete b;
  return CUOT_OK;
}
