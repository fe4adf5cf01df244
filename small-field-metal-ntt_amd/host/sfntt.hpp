// sfntt.hpp -- header-only C++17 mirror of the reference's operator API over
// the C ABI in include/sfntt.h.
//
// Names, argument meaning and error behaviour follow the reference's Rust
// backends, so C++ callers read like the reference's own call sites:
//   NttBackend::{name, forward_ntt, inverse_ntt, pointwise_mul}   (src/ntt/mod.rs:26-49)
//   <Backend>::{forward_ntt_gpu, inverse_ntt_gpu}(&[u32], log_n) -> (Vec<u32>, ns)
//                                                                 (src/ntt/metal_r2.rs:69-173)
//   BbBatchNtt::{forward,inverse}_ntt_batch_gpu(&[u32], log_n, batch)  (src/ntt/babybear/batch.rs:76-180)
//   CosetLdeBatch::execute(&[u32], log_n, batch, added_bits) -> LdeResult  (src/lde/mod.rs:43-107)
//   TwiddleCache<F>                                              (src/ntt/twiddle_cache.rs:17-82)
//   NttError variants as exception types                         (src/ntt/mod.rs:51-69)
// The reference's Metal type names are kept as aliases (MetalCtGsR2, ...).
//
// A Context wraps one sfn_ctx and is externally synchronised, like the
// reference's single-threaded MetalContext (src/gpu.rs:26-417).  Backends
// share a Context through std::shared_ptr.  Link with -lsfntt.
#pragma once

#include <sfntt.h>

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace sfntt {

// ------------------------------------------------------------------ errors
class NttError : public std::runtime_error {
 public:
  NttError(sfn_status status, const std::string &msg) : std::runtime_error(msg), status_(status) {}
  sfn_status status() const { return status_; }

 private:
  sfn_status status_;
};

#define SFNTT_ERROR(Name)                                                           \
  class Name : public NttError {                                                    \
   public:                                                                          \
    Name(sfn_status s, const std::string &m) : NttError(s, #Name ": " + m) {}      \
  };
SFNTT_ERROR(InvalidSize)
SFNTT_ERROR(DeviceNotFound)
SFNTT_ERROR(BufferAllocFailed)
SFNTT_ERROR(ShaderCompileError)
SFNTT_ERROR(GpuExecutionError)
SFNTT_ERROR(CorrectnessMismatch)
SFNTT_ERROR(Unsupported)
SFNTT_ERROR(InvalidArgument)
SFNTT_ERROR(Busy)
#undef SFNTT_ERROR

// Throws the exception type that matches a non-OK status.
inline void check(sfn_status s, const sfn_ctx *ctx = nullptr) {
  if (s == SFN_OK) return;
  std::string msg = sfn_status_string(s);
  if (ctx) {
    const char *d = sfn_last_error(ctx);
    if (d && *d) msg += std::string(": ") + d;
  }
  switch (s) {
    case SFN_ERR_INVALID_SIZE: throw InvalidSize(s, msg);
    case SFN_ERR_DEVICE_NOT_FOUND: throw DeviceNotFound(s, msg);
    case SFN_ERR_BUFFER_ALLOC_FAILED: throw BufferAllocFailed(s, msg);
    case SFN_ERR_KERNEL_LOAD: throw ShaderCompileError(s, msg);
    case SFN_ERR_GPU_EXECUTION: throw GpuExecutionError(s, msg);
    case SFN_ERR_CORRECTNESS: throw CorrectnessMismatch(s, msg);
    case SFN_ERR_UNSUPPORTED: throw Unsupported(s, msg);
    case SFN_ERR_INVALID_ARGUMENT: throw InvalidArgument(s, msg);
    case SFN_ERR_BUSY: throw Busy(s, msg);
    default: throw NttError(s, msg);
  }
}

enum class Field : int { M31 = SFN_FIELD_M31, BabyBear = SFN_FIELD_BABYBEAR, KoalaBear = SFN_FIELD_KOALABEAR };
enum class Variant : int {
  CtDitR2 = SFN_CT_DIT_R2,
  CtGsR2 = SFN_CT_GS_R2,
  StockhamR2 = SFN_STOCKHAM_R2,
  CtGsR4 = SFN_CT_GS_R4,
  Auto = SFN_VARIANT_AUTO
};
enum class Layout : int { SoA = SFN_LAYOUT_SOA, RowMajor = SFN_LAYOUT_ROW_MAJOR };

// --------------------------------------------------------------- context
class Context {
 public:
  explicit Context(int device = 0) { check(sfn_ctx_create(device, &ctx_)); }
  ~Context() { sfn_ctx_destroy(ctx_); }
  Context(const Context &) = delete;
  Context &operator=(const Context &) = delete;

  sfn_ctx *raw() const { return ctx_; }
  void check(sfn_status s) const { sfntt::check(s, ctx_); }

  sfn_device_info device_info() const {
    sfn_device_info info{};
    check(sfn_device_info_get(ctx_, &info));
    return info;
  }

  // Host-pointer transforms (upload, launches, download); returns device ns.
  uint64_t ntt(Field f, Variant v, sfn_direction dir, uint32_t log_n, uint32_t batch, Layout layout,
               const uint32_t *in, uint32_t *out) const {
    uint64_t ns = 0;
    check(sfn_ntt(ctx_, (sfn_field)f, (sfn_variant)v, dir, log_n, batch, (sfn_layout)layout, in, out, &ns));
    return ns;
  }
  uint64_t coset_lde(uint32_t log_n, uint32_t batch, uint32_t added_bits, uint32_t shift_raw, Layout layout,
                     const uint32_t *in, uint32_t *out, bool natural = false) const {
    uint64_t ns = 0;
    check((natural ? sfn_coset_lde_natural : sfn_coset_lde_batch)(ctx_, log_n, batch, added_bits, shift_raw,
                                                                   (sfn_layout)layout, in, out, &ns));
    return ns;
  }

 private:
  sfn_ctx *ctx_ = nullptr;
};

inline std::shared_ptr<Context> default_context(int device = 0) {
  return std::make_shared<Context>(device);
}

// Power-of-two length -> log2 (InvalidSize otherwise), as the backends check.
inline uint32_t log2_exact(size_t n) {
  if (n == 0 || (n & (n - 1))) throw InvalidSize(SFN_ERR_INVALID_SIZE, std::to_string(n));
  uint32_t l = 0;
  while ((size_t(1) << l) < n) ++l;
  return l;
}

// (output, device ns): the `*_gpu` return pair.
using GpuResult = std::pair<std::vector<uint32_t>, uint64_t>;

// -------------------------------------------------------------- backends
// One NttBackend<F> variant.  forward_ntt / inverse_ntt work in place on raw
// field values (M31 canonical, BabyBear Montgomery raw); the twiddle argument
// of the reference's trait is ignored there too, so it is omitted here.
class Backend {
 public:
  Backend(std::shared_ptr<Context> ctx, Field field, Variant variant, std::string name)
      : ctx_(std::move(ctx)), field_(field), variant_(variant), name_(std::move(name)) {}

  const std::string &name() const { return name_; }
  Field field() const { return field_; }

  GpuResult forward_ntt_gpu(const std::vector<uint32_t> &input, uint32_t log_n) const {
    return gpu(SFN_FORWARD, input, log_n, 1, Layout::SoA);
  }
  GpuResult inverse_ntt_gpu(const std::vector<uint32_t> &input, uint32_t log_n) const {
    return gpu(SFN_INVERSE, input, log_n, 1, Layout::SoA);
  }

  void forward_ntt(std::vector<uint32_t> &data) const {
    const uint32_t log_n = log2_exact(data.size());
    if (log_n) data = forward_ntt_gpu(data, log_n).first;
  }
  void inverse_ntt(std::vector<uint32_t> &data) const {
    const uint32_t log_n = log2_exact(data.size());
    if (log_n) data = inverse_ntt_gpu(data, log_n).first;
  }
  void pointwise_mul(const std::vector<uint32_t> &a, const std::vector<uint32_t> &b,
                     std::vector<uint32_t> &out) const {
    if (a.size() != b.size() || a.size() != out.size())
      throw InvalidSize(SFN_ERR_INVALID_SIZE, std::to_string(a.size()));
    ctx_->check(sfn_pointwise_mul(ctx_->raw(), (sfn_field)field_, a.data(), b.data(), out.data(), a.size()));
  }

 protected:
  GpuResult gpu(sfn_direction dir, const std::vector<uint32_t> &input, uint32_t log_n, uint32_t batch,
                Layout layout) const {
    const uint32_t max_log = field_ == Field::M31 ? 30u : (field_ == Field::KoalaBear ? 24u : 27u);
    if (batch == 0 || log_n > max_log || input.size() != (size_t(batch) << log_n))
      throw InvalidSize(SFN_ERR_INVALID_SIZE, std::to_string(input.size()));
    if (log_n == 0) return {input, 0};
    std::vector<uint32_t> out(input.size());
    const uint64_t ns = ctx_->ntt(field_, variant_, dir, log_n, batch, layout, input.data(), out.data());
    return {std::move(out), ns};
  }

  std::shared_ptr<Context> ctx_;
  Field field_;
  Variant variant_;
  std::string name_;
};

// Batched BabyBear NTT, SoA data[col * n + row] (row-major optional).
class BbBatchNtt : public Backend {
 public:
  explicit BbBatchNtt(std::shared_ptr<Context> ctx)
      : Backend(std::move(ctx), Field::BabyBear, Variant::Auto, "bb-hip-batch") {}
  BbBatchNtt(std::shared_ptr<Context> ctx, Field f, std::string name)
      : Backend(std::move(ctx), f, Variant::Auto, std::move(name)) {}
  GpuResult forward_ntt_batch_gpu(const std::vector<uint32_t> &data, uint32_t log_n, uint32_t batch,
                                  Layout layout = Layout::SoA) const {
    return gpu(SFN_FORWARD, data, log_n, batch, layout);
  }
  GpuResult inverse_ntt_batch_gpu(const std::vector<uint32_t> &data, uint32_t log_n, uint32_t batch,
                                  Layout layout = Layout::SoA) const {
    return gpu(SFN_INVERSE, data, log_n, batch, layout);
  }
};

// M31 circle NTT variants (src/ntt/m31/*.rs) and BabyBear variants
// (src/ntt/babybear/*.rs) under the reference's type names.
inline Backend MetalCtDitR2(std::shared_ptr<Context> c) { return Backend(std::move(c), Field::M31, Variant::CtDitR2, "hip-ct-dit-r2"); }
inline Backend MetalCtGsR2(std::shared_ptr<Context> c) { return Backend(std::move(c), Field::M31, Variant::CtGsR2, "hip-ct-gs-r2"); }
inline Backend MetalStockhamR2(std::shared_ptr<Context> c) { return Backend(std::move(c), Field::M31, Variant::StockhamR2, "hip-stockham-r2"); }
inline Backend MetalCtGsR4(std::shared_ptr<Context> c) { return Backend(std::move(c), Field::M31, Variant::CtGsR4, "hip-ct-gs-r4"); }
inline Backend BbMetalCtDitR2(std::shared_ptr<Context> c) { return Backend(std::move(c), Field::BabyBear, Variant::CtDitR2, "bb-hip-ct-dit-r2"); }
inline Backend BbMetalR2(std::shared_ptr<Context> c) { return Backend(std::move(c), Field::BabyBear, Variant::CtGsR2, "bb-hip-r2"); }
inline Backend BbMetalStockhamR2(std::shared_ptr<Context> c) { return Backend(std::move(c), Field::BabyBear, Variant::StockhamR2, "bb-hip-stockham-r2"); }
inline Backend BbMetalCtGsR4(std::shared_ptr<Context> c) { return Backend(std::move(c), Field::BabyBear, Variant::CtGsR4, "bb-hip-ct-gs-r4"); }

// KoalaBear (README.md:20, planned in the reference): the BabyBear variants'
// conventions with p = 2^31 - 2^24 + 1 and two-adicity 24.
inline Backend KbMetalCtDitR2(std::shared_ptr<Context> c) { return Backend(std::move(c), Field::KoalaBear, Variant::CtDitR2, "kb-hip-ct-dit-r2"); }
inline Backend KbMetalR2(std::shared_ptr<Context> c) { return Backend(std::move(c), Field::KoalaBear, Variant::CtGsR2, "kb-hip-r2"); }
inline Backend KbMetalStockhamR2(std::shared_ptr<Context> c) { return Backend(std::move(c), Field::KoalaBear, Variant::StockhamR2, "kb-hip-stockham-r2"); }
inline Backend KbMetalCtGsR4(std::shared_ptr<Context> c) { return Backend(std::move(c), Field::KoalaBear, Variant::CtGsR4, "kb-hip-ct-gs-r4"); }
class KbBatchNtt : public BbBatchNtt {
 public:
  explicit KbBatchNtt(std::shared_ptr<Context> ctx) : BbBatchNtt(std::move(ctx), Field::KoalaBear, "kb-hip-batch") {}
};

// Device-resident twiddle tables keyed by (field, log_n); the reference's
// TwiddleCache keeps only the last log_n on the host.
class TwiddleCache {
 public:
  explicit TwiddleCache(std::shared_ptr<Context> ctx) : ctx_(std::move(ctx)) {}
  void prepare(Field f, uint32_t log_n) const { ctx_->check(sfn_twiddles_prepare(ctx_->raw(), (sfn_field)f, log_n)); }
  size_t bytes() const {
    size_t b = 0;
    ctx_->check(sfn_twiddles_bytes(ctx_->raw(), &b));
    return b;
  }

 private:
  std::shared_ptr<Context> ctx_;
};

// ------------------------------------------------------------ coset LDE
struct LdeResult {
  std::vector<uint32_t> data;  // SoA data[col * (n << added_bits) + row] (or row-major)
  uint64_t total_ns = 0;
};

class CosetLdeBatch {
 public:
  explicit CosetLdeBatch(std::shared_ptr<Context> ctx) : ctx_(std::move(ctx)) {}
  // shift_raw = 0: the reference shift two_adic_generator(log_n + added_bits + 1).
  LdeResult execute(const std::vector<uint32_t> &data, uint32_t log_n, uint32_t batch, uint32_t added_bits,
                    uint32_t shift_raw = 0, Layout layout = Layout::SoA) const {
    return run(data, log_n, batch, added_bits, shift_raw, layout, false);
  }
  // Natural row order on both sides (input row i = P(g_n^i), output row k =
  // P(shift * g_N^k)), the order of Plonky3's coset_lde_batch.
  LdeResult execute_natural(const std::vector<uint32_t> &data, uint32_t log_n, uint32_t batch, uint32_t added_bits,
                            uint32_t shift_raw = 0, Layout layout = Layout::SoA) const {
    return run(data, log_n, batch, added_bits, shift_raw, layout, true);
  }
  sfn_device_info device_info() const { return ctx_->device_info(); }

 private:
  LdeResult run(const std::vector<uint32_t> &data, uint32_t log_n, uint32_t batch, uint32_t added_bits,
                uint32_t shift_raw, Layout layout, bool natural) const {
    if (data.size() != (size_t(batch) << log_n))  // src/lde/mod.rs:54-56
      throw InvalidSize(SFN_ERR_INVALID_SIZE, std::to_string(data.size()));
    if (log_n + added_bits + 1 > 27)  // two_adic_generator bound (src/field/babybear.rs:69-76)
      throw InvalidSize(SFN_ERR_INVALID_SIZE, "log_n + added_bits + 1 > 27");
    LdeResult r;
    r.data.resize(size_t(batch) << (log_n + added_bits));
    r.total_ns = ctx_->coset_lde(log_n, batch, added_bits, shift_raw, layout, data.data(), r.data.data(), natural);
    return r;
  }

  std::shared_ptr<Context> ctx_;
};

}  // namespace sfntt
