// sigops.hpp -- header-only C++ mirror of the reference's public API over the C ABI (sigops.h).
//
// Same names, argument meaning and results as the reference's async Rust functions:
//   secp256k1_ecdsa::ecrecover(signatures, messages, table_limbs, log_limb_size)   src/secp256k1_ecdsa.rs:61-66
//   secp256k1_ecdsa::ecrecover_single_shader(signatures, messages, log_limb_size)  src/secp256k1_ecdsa.rs:215-219
//   secp256r1_ecdsa::ecrecover / ecrecover_single_shader                         src/secp256r1_ecdsa.rs:62-67, :216-220
//   ed25519_eddsa::ecverify(signatures, messages, verifying_keys, table_limbs, log_limb_size)
//   ed25519_eddsa::ecverify_single(signatures, messages, verifying_keys, log_limb_size)
//                                                                                src/ed25519_eddsa.rs:67-73, :259-264
// and, next to the path: secp256k1_ecdsa::verify / secp256r1_ecdsa::verify (known public key) and
// ed25519_eddsa::ecverify_strict (dalek verify_strict)
// Rust's Result<_, ShaderFailureError> (src/lib.rs:12-14) becomes a thrown sigops::ShaderFailureError;
// the reference's panics on bad arguments (length mismatch, src/secp256k1_ecdsa.rs:21) become
// std::invalid_argument. A table argument is converted on the device and used as the fixed base, as
// the reference uses it (sigops_*_table); one that is not [P, 2P, ..., 16P] throws std::runtime_error.
#pragma once
#include <array>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "sigops.h"

namespace sigops {

using Signature = std::array<uint8_t, 64>;  // fuel Signature / Bytes64: r || s, parity in bit 7 of byte 32
using Bytes64 = std::array<uint8_t, 64>;
using Message = std::array<uint8_t, 32>;    // fuel Message: a 32-byte digest
using VerifyingKey = std::array<uint8_t, 32>;
using Ed25519Signature = std::array<uint8_t, 64>;  // R || s

struct ShaderFailureError : std::runtime_error {
  explicit ShaderFailureError(const std::string& w) : std::runtime_error(w) {}
};

class Context {
 public:
  explicit Context(const sigops_config* cfg = nullptr) : c_(sigops_create(cfg)) {
    if (!c_) throw std::runtime_error("sigops_create failed");
  }
  ~Context() { sigops_destroy(c_); }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  sigops_ctx* get() const { return c_; }
  void check(int rc) const {
    if (rc == SIGOPS_OK) return;
    const std::string msg = sigops_last_error(c_);
    if (rc == SIGOPS_E_CANARY) throw ShaderFailureError(msg);
    throw std::runtime_error("libsigops error " + std::to_string(rc) + ": " + msg);
  }

 private:
  sigops_ctx* c_;
};

inline Context& default_context() {
  static Context ctx;
  return ctx;
}

namespace detail {
inline void check_log_limb_size(uint32_t l) {
  if (l < 11 || l > 15) throw std::invalid_argument("log_limb_size must be in 11..15");
}
inline void check_table(const std::vector<uint32_t>& t, size_t coords, uint32_t l) {
  check_log_limb_size(l);
  if (t.size() != 16 * coords * (256 / l + 1)) throw std::invalid_argument("table_limbs has the wrong size");
}
// table: NULL (the resident generator table) or the caller's table_limbs in the reference's layout
template <class Sig>
std::vector<std::vector<uint8_t>> recover(int curve, const std::vector<Sig>& signatures,
                                          const std::vector<Message>& messages, Context& ctx,
                                          const std::vector<uint32_t>* table = nullptr, uint32_t log_limb_size = 0) {
  if (signatures.size() != messages.size()) throw std::invalid_argument("signatures and messages differ in length");
  const size_t n = signatures.size();
  if (n > (size_t(1) << 30)) throw std::invalid_argument("at most 2^30 signatures");  // init's assert, :21-22
  if (n == 0) return {};
  std::vector<uint8_t> sigs(n * 64), digs(n * 32), out(n * 64);
  for (size_t i = 0; i < n; i++) {
    memcpy(&sigs[i * 64], signatures[i].data(), 64);
    memcpy(&digs[i * 32], messages[i].data(), 32);
  }
  if (table)
    ctx.check((curve == 0 ? sigops_secp256k1_recover_table : sigops_secp256r1_recover_table)(
        ctx.get(), sigs.data(), digs.data(), n, table->data(), table->size(), log_limb_size, out.data(), nullptr));
  else
    ctx.check(curve == 0 ? sigops_secp256k1_recover(ctx.get(), sigs.data(), digs.data(), n, out.data(), nullptr)
                         : sigops_secp256r1_recover(ctx.get(), sigs.data(), digs.data(), n, out.data(), nullptr));
  std::vector<std::vector<uint8_t>> res(n);
  for (size_t i = 0; i < n; i++) res[i].assign(out.begin() + i * 64, out.begin() + i * 64 + 64);
  return res;
}
// fuel Signature::verify(&public_key, &message) per item (public keys: 64 bytes X || Y)
template <class Sig>
std::vector<bool> verify(int curve, const std::vector<Sig>& signatures, const std::vector<Message>& messages,
                         const std::vector<Bytes64>& public_keys, Context& ctx) {
  if (signatures.size() != messages.size() || signatures.size() != public_keys.size())
    throw std::invalid_argument("signatures, messages and public keys differ in length");
  const size_t n = signatures.size();
  if (n > (size_t(1) << 30)) throw std::invalid_argument("at most 2^30 signatures");
  if (n == 0) return {};
  std::vector<uint8_t> sigs(n * 64), digs(n * 32), pubs(n * 64), valid(n);
  for (size_t i = 0; i < n; i++) {
    memcpy(&sigs[i * 64], signatures[i].data(), 64);
    memcpy(&digs[i * 32], messages[i].data(), 32);
    memcpy(&pubs[i * 64], public_keys[i].data(), 64);
  }
  ctx.check(curve == 0 ? sigops_secp256k1_verify(ctx.get(), sigs.data(), digs.data(), pubs.data(), n, valid.data(), nullptr)
                       : sigops_secp256r1_verify(ctx.get(), sigs.data(), digs.data(), pubs.data(), n, valid.data(), nullptr));
  std::vector<bool> res(n);
  for (size_t i = 0; i < n; i++) res[i] = valid[i] == 1;
  return res;
}
}  // namespace detail

namespace secp256k1_ecdsa {
// verification with a known key (libsecp256k1 rules: lower-S only); no counterpart in the reference's API
inline std::vector<bool> verify(const std::vector<Signature>& signatures, const std::vector<Message>& messages,
                                const std::vector<Bytes64>& public_keys, Context& ctx = default_context()) {
  return detail::verify(0, signatures, messages, public_keys, ctx);
}
inline std::vector<std::vector<uint8_t>> ecrecover(const std::vector<Signature>& signatures,
                                                   const std::vector<Message>& messages,
                                                   const std::vector<uint32_t>& table_limbs, uint32_t log_limb_size,
                                                   Context& ctx = default_context()) {
  detail::check_table(table_limbs, 2, log_limb_size);
  return detail::recover(0, signatures, messages, ctx, &table_limbs, log_limb_size);
}
inline std::vector<std::vector<uint8_t>> ecrecover_single_shader(const std::vector<Signature>& signatures,
                                                                 const std::vector<Message>& messages,
                                                                 uint32_t log_limb_size,
                                                                 Context& ctx = default_context()) {
  detail::check_log_limb_size(log_limb_size);
  return detail::recover(0, signatures, messages, ctx);
}
inline std::vector<std::vector<uint8_t>> ecrecover_single(const std::vector<Signature>& signatures,
                                                          const std::vector<Message>& messages, uint32_t log_limb_size,
                                                          Context& ctx = default_context()) {
  return ecrecover_single_shader(signatures, messages, log_limb_size, ctx);
}
}  // namespace secp256k1_ecdsa

namespace secp256r1_ecdsa {
// verification with a known key (p256 / ecdsa 0.16 rules); no counterpart in the reference's API
inline std::vector<bool> verify(const std::vector<Bytes64>& signatures, const std::vector<Message>& messages,
                                const std::vector<Bytes64>& public_keys, Context& ctx = default_context()) {
  return detail::verify(1, signatures, messages, public_keys, ctx);
}
inline std::vector<std::vector<uint8_t>> ecrecover(const std::vector<Bytes64>& signatures,
                                                   const std::vector<Message>& messages,
                                                   const std::vector<uint32_t>& table_limbs, uint32_t log_limb_size,
                                                   Context& ctx = default_context()) {
  detail::check_table(table_limbs, 2, log_limb_size);
  return detail::recover(1, signatures, messages, ctx, &table_limbs, log_limb_size);
}
inline std::vector<std::vector<uint8_t>> ecrecover_single_shader(const std::vector<Bytes64>& signatures,
                                                                 const std::vector<Message>& messages,
                                                                 uint32_t log_limb_size,
                                                                 Context& ctx = default_context()) {
  detail::check_log_limb_size(log_limb_size);
  return detail::recover(1, signatures, messages, ctx);
}
inline std::vector<std::vector<uint8_t>> ecrecover_single(const std::vector<Bytes64>& signatures,
                                                          const std::vector<Message>& messages, uint32_t log_limb_size,
                                                          Context& ctx = default_context()) {
  return ecrecover_single_shader(signatures, messages, log_limb_size, ctx);
}
}  // namespace secp256r1_ecdsa

namespace ed25519_eddsa {
namespace detail_ed {
inline std::vector<bool> verify(const std::vector<Ed25519Signature>& signatures, const std::vector<Message>& messages,
                                const std::vector<VerifyingKey>& verifying_keys, Context& ctx,
                                const std::vector<uint32_t>* table = nullptr, uint32_t log_limb_size = 0) {
  if (signatures.size() != messages.size()) throw std::invalid_argument("signatures and messages differ in length");
  const size_t n = signatures.size();
  if (n == 0) return {};
  if (verifying_keys.size() < n) throw std::out_of_range("fewer verifying keys than signatures");
  std::vector<uint8_t> sigs(n * 64), pks(n * 32), msgs(n * 32), valid(n);
  std::vector<uint64_t> off(n + 1);
  for (size_t i = 0; i < n; i++) {
    memcpy(&sigs[i * 64], signatures[i].data(), 64);
    memcpy(&pks[i * 32], verifying_keys[i].data(), 32);
    memcpy(&msgs[i * 32], messages[i].data(), 32);
    off[i] = 32 * i;
  }
  off[n] = 32 * n;
  if (table)
    ctx.check(sigops_ed25519_verify_table(ctx.get(), sigs.data(), pks.data(), msgs.data(), off.data(), n,
                                          table->data(), table->size(), log_limb_size, valid.data(), nullptr));
  else
    ctx.check(
        sigops_ed25519_verify(ctx.get(), sigs.data(), pks.data(), msgs.data(), off.data(), n, valid.data(), nullptr));
  std::vector<bool> res(n);
  for (size_t i = 0; i < n; i++) res[i] = valid[i] == 1;
  return res;
}
}  // namespace detail_ed

inline std::vector<bool> ecverify(const std::vector<Ed25519Signature>& signatures, const std::vector<Message>& messages,
                                  const std::vector<VerifyingKey>& verifying_keys,
                                  const std::vector<uint32_t>& table_limbs, uint32_t log_limb_size,
                                  Context& ctx = default_context()) {
  detail::check_table(table_limbs, 3, log_limb_size);
  return detail_ed::verify(signatures, messages, verifying_keys, ctx, &table_limbs, log_limb_size);
}
inline std::vector<bool> ecverify_single(const std::vector<Ed25519Signature>& signatures,
                                         const std::vector<Message>& messages,
                                         const std::vector<VerifyingKey>& verifying_keys, uint32_t log_limb_size,
                                         Context& ctx = default_context()) {
  detail::check_log_limb_size(log_limb_size);
  return detail_ed::verify(signatures, messages, verifying_keys, ctx);
}
namespace detail_ed {
inline std::vector<bool> verify_bytes(const std::vector<Ed25519Signature>& signatures,
                                      const std::vector<std::vector<uint8_t>>& messages,
                                      const std::vector<VerifyingKey>& verifying_keys, Context& ctx, bool strict) {
  if (signatures.size() != messages.size()) throw std::invalid_argument("signatures and messages differ in length");
  const size_t n = signatures.size();
  if (n == 0) return {};
  if (verifying_keys.size() < n) throw std::out_of_range("fewer verifying keys than signatures");
  std::vector<uint8_t> sigs(n * 64), pks(n * 32), valid(n), msgs;
  std::vector<uint64_t> off(n + 1);
  for (size_t i = 0; i < n; i++) {
    memcpy(&sigs[i * 64], signatures[i].data(), 64);
    memcpy(&pks[i * 32], verifying_keys[i].data(), 32);
    off[i] = msgs.size();
    msgs.insert(msgs.end(), messages[i].begin(), messages[i].end());
  }
  off[n] = msgs.size();
  ctx.check((strict ? sigops_ed25519_verify_strict : sigops_ed25519_verify)(
      ctx.get(), sigs.data(), pks.data(), msgs.empty() ? nullptr : msgs.data(), off.data(), n, valid.data(), nullptr));
  std::vector<bool> res(n);
  for (size_t i = 0; i < n; i++) res[i] = valid[i] == 1;
  return res;
}
}  // namespace detail_ed

// variable-length extension: message i is messages[i] (any length), hashed as given; the
// reference's Message is fixed at 32 bytes (src/ed25519_eddsa.rs:39)
inline std::vector<bool> ecverify_bytes(const std::vector<Ed25519Signature>& signatures,
                                        const std::vector<std::vector<uint8_t>>& messages,
                                        const std::vector<VerifyingKey>& verifying_keys,
                                        Context& ctx = default_context()) {
  return detail_ed::verify_bytes(signatures, messages, verifying_keys, ctx, false);
}
// ed25519-dalek verify_strict rules (src/curve_algos/ed25519_eddsa.rs:197-198 names it)
inline std::vector<bool> ecverify_strict(const std::vector<Ed25519Signature>& signatures,
                                         const std::vector<std::vector<uint8_t>>& messages,
                                         const std::vector<VerifyingKey>& verifying_keys,
                                         Context& ctx = default_context()) {
  return detail_ed::verify_bytes(signatures, messages, verifying_keys, ctx, true);
}
}  // namespace ed25519_eddsa

}  // namespace sigops
