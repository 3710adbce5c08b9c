// ref_aes.cpp -- C wrapper around the reference's vendored cryptoTools AES
// (silent/lib/cryptoTools/Crypto/AES.{h,cpp}), compiled from the reference
// sources where they lie.  Test infrastructure only: used to cross-check the
// oracle's AES-128.
#include <cstring>

#include "cryptoTools/Crypto/AES.h"

extern "C" void ref_aes_ecb(const void* key16, const void* in, void* out, size_t n) {
  osuCrypto::block k;
  std::memcpy(&k, key16, 16);
  osuCrypto::AES aes;
  aes.setKey(k);
  aes.ecbEncBlocks(static_cast<const osuCrypto::block*>(in), n, static_cast<osuCrypto::block*>(out));
}
