// Tests of the C++ host mirror (small-field-metal-ntt_amd/host/sfntt.hpp).
//
//   host_api_test --no-device   error mapping and size checks; expects that no
//                               GPU is visible (Context() -> DeviceNotFound)
//   host_api_test --parity      on a GPU: every backend type against the CPU
//                               oracle (linked from oracle/liboracle.so) on
//                               seeded inputs, bit for bit
// Prints one line per check and exits non-zero on the first failure.
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "sfntt.hpp"
#include "sfntt_oracle.h"

namespace {

int failures = 0;

void expect(bool ok, const std::string &what) {
  std::printf("%s %s\n", ok ? "ok  " : "FAIL", what.c_str());
  std::fflush(stdout);
  if (!ok) ++failures;
}

template <class E>
bool throws(const std::function<void()> &f) {
  try {
    f();
  } catch (const E &) {
    return true;
  } catch (...) {
    return false;
  }
  return false;
}

std::vector<uint32_t> lcg(uint64_t seed, size_t n, bool babybear) {
  std::vector<uint32_t> v(n);
  orc_lcg_fill(seed, babybear ? 0x78000001u : 0x7FFFFFFFu, babybear ? 1 : 0, v.data(), n);
  return v;
}

std::vector<uint32_t> lcg_kb(uint64_t seed, size_t n) {
  std::vector<uint32_t> v(n);
  orc_lcg_fill(seed, ORC_KB_P, 2, v.data(), n);
  return v;
}

int no_device() {
  using namespace sfntt;
  expect(throws<InvalidSize>([] { check(SFN_ERR_INVALID_SIZE); }), "check(INVALID_SIZE) throws InvalidSize");
  expect(throws<DeviceNotFound>([] { check(SFN_ERR_DEVICE_NOT_FOUND); }), "check(DEVICE_NOT_FOUND) throws DeviceNotFound");
  expect(throws<ShaderCompileError>([] { check(SFN_ERR_KERNEL_LOAD); }), "check(KERNEL_LOAD) throws ShaderCompileError");
  expect(throws<Busy>([] { check(SFN_ERR_BUSY); }), "check(BUSY) throws Busy");
  expect(throws<NttError>([] { check(SFN_ERR_UNSUPPORTED); }), "every error derives from NttError");
  bool no_throw = true;
  try {
    check(SFN_OK);
  } catch (...) {
    no_throw = false;
  }
  expect(no_throw, "check(OK) is silent");
  expect(log2_exact(1) == 0 && log2_exact(1u << 20) == 20, "log2_exact of powers of two");
  expect(throws<InvalidSize>([] { log2_exact(3); }) && throws<InvalidSize>([] { log2_exact(0); }),
         "log2_exact rejects 0 and non-powers of two");
  bool device = false;
  try {
    Context c(0);
    device = true;
  } catch (const DeviceNotFound &) {
  }
  if (device) std::printf("skip Context() without a GPU (a gfx950 device is visible)\n");
  else expect(true, "Context() without a GPU throws DeviceNotFound");
  return failures ? 1 : 0;
}

int parity() {
  using namespace sfntt;
  auto ctx = default_context(0);
  const sfn_device_info info = ctx->device_info();
  expect(std::strncmp(info.gcn_arch, "gfx950", 6) == 0, std::string("device ") + info.gcn_arch);

  // BabyBear variants, forward and inverse, 2^4 .. 2^16
  std::vector<Backend> bb = {BbMetalCtDitR2(ctx), BbMetalR2(ctx), BbMetalStockhamR2(ctx), BbMetalCtGsR4(ctx)};
  for (uint32_t log_n : {4u, 11u, 16u}) {
    const size_t n = size_t(1) << log_n;
    const std::vector<uint32_t> x = lcg(12345 + n, n, true);
    std::vector<uint32_t> want = x;
    orc_bb_forward(want.data(), n);
    for (const Backend &b : bb) {
      std::vector<uint32_t> y = x;
      b.forward_ntt(y);
      expect(y == want, b.name() + " forward 2^" + std::to_string(log_n));
      if (b.name() == "bb-hip-ct-dit-r2") continue;  // V1 is forward only in the reference
      b.inverse_ntt(y);
      expect(y == x, b.name() + " inverse(forward) 2^" + std::to_string(log_n));
    }
  }
  // M31 variants
  std::vector<Backend> m31 = {MetalCtDitR2(ctx), MetalCtGsR2(ctx), MetalStockhamR2(ctx), MetalCtGsR4(ctx)};
  for (uint32_t log_n : {3u, 12u, 18u}) {
    const size_t n = size_t(1) << log_n;
    const std::vector<uint32_t> x = lcg(12345 + n, n, false);
    std::vector<uint32_t> want = x;
    orc_m31_forward(want.data(), n);
    for (const Backend &b : m31) {
      GpuResult r = b.forward_ntt_gpu(x, log_n);
      expect(r.first == want, b.name() + " forward_ntt_gpu 2^" + std::to_string(log_n));
    }
  }
  // KoalaBear variants and batch
  std::vector<Backend> kb = {KbMetalCtDitR2(ctx), KbMetalR2(ctx), KbMetalStockhamR2(ctx), KbMetalCtGsR4(ctx)};
  for (uint32_t log_n : {5u, 13u, 17u}) {
    const size_t n = size_t(1) << log_n;
    const std::vector<uint32_t> x = lcg_kb(777 + n, n);
    std::vector<uint32_t> want = x;
    orc_kb_forward(want.data(), n);
    for (const Backend &b : kb) {
      std::vector<uint32_t> y = x;
      b.forward_ntt(y);
      expect(y == want, b.name() + " forward 2^" + std::to_string(log_n));
      b.inverse_ntt(y);
      expect(y == x, b.name() + " inverse(forward) 2^" + std::to_string(log_n));
    }
  }
  {
    KbBatchNtt batch(ctx);
    const std::vector<uint32_t> x = lcg_kb(31, size_t(5) << 11);
    std::vector<uint32_t> want = x;
    orc_ntt_batch(2, 1, 11, 5, 0, want.data(), 1);
    expect(batch.inverse_ntt_batch_gpu(x, 11, 5).first == want, "KbBatchNtt inverse SoA");
    expect(throws<InvalidSize>([&] { batch.forward_ntt_batch_gpu(std::vector<uint32_t>(size_t(1) << 25), 25, 1); }),
           "KoalaBear beyond two-adicity 24 throws InvalidSize");
  }
  // batched BabyBear, SoA and row-major
  {
    BbBatchNtt batch(ctx);
    const uint32_t log_n = 12, cols = 6;
    const std::vector<uint32_t> x = lcg(99, size_t(cols) << log_n, true);
    for (Layout lay : {Layout::SoA, Layout::RowMajor}) {
      std::vector<uint32_t> want = x;
      orc_ntt_batch(1, 0, log_n, cols, (int)lay, want.data(), 1);
      GpuResult r = batch.forward_ntt_batch_gpu(x, log_n, cols, lay);
      expect(r.first == want, std::string("BbBatchNtt forward ") + (lay == Layout::SoA ? "SoA" : "row-major"));
    }
  }
  // coset LDE, reference shift and the field generator
  {
    CosetLdeBatch lde(ctx);
    const uint32_t log_n = 10, cols = 5;
    const std::vector<uint32_t> x = lcg(42 + log_n, size_t(cols) << log_n, true);
    for (uint32_t shift : {0u, orc_bb_to_monty(31)}) {
      for (uint32_t a : {1u, 2u}) {
        std::vector<uint32_t> want(size_t(cols) << (log_n + a));
        orc_coset_lde_batch(x.data(), log_n, cols, a, shift, 0, want.data(), 1);
        LdeResult r = lde.execute(x, log_n, cols, a, shift);
        expect(r.data == want, "CosetLdeBatch.execute 2^10 x 5, added_bits " + std::to_string(a) +
                                   (shift ? ", shift 31" : ", reference shift"));
      }
    }
  }
  // natural row order on both sides: bit-reversed rows around the reference LDE
  {
    CosetLdeBatch lde(ctx);
    const uint32_t log_n = 9, cols = 3, a = 2;
    const size_t n = size_t(1) << log_n, N = n << a;
    const std::vector<uint32_t> x = lcg(77, cols * n, true);
    auto brev = [](size_t v, uint32_t bits) {
      size_t r = 0;
      for (uint32_t b = 0; b < bits; ++b) r |= ((v >> b) & 1) << (bits - 1 - b);
      return r;
    };
    std::vector<uint32_t> xr(x.size()), y(cols * N), want(cols * N);
    for (size_t c = 0; c < cols; ++c)
      for (size_t r = 0; r < n; ++r) xr[c * n + r] = x[c * n + brev(r, log_n)];
    orc_coset_lde_batch(xr.data(), log_n, cols, a, orc_bb_to_monty(31), 0, y.data(), 1);
    for (size_t c = 0; c < cols; ++c)
      for (size_t r = 0; r < N; ++r) want[c * N + r] = y[c * N + brev(r, log_n + a)];
    LdeResult got = lde.execute_natural(x, log_n, cols, a, orc_bb_to_monty(31));
    expect(got.data == want, "CosetLdeBatch.execute_natural 2^9 x 3, added_bits 2, shift 31");
  }
  // pointwise_mul
  {
    Backend b = BbMetalR2(ctx);
    const std::vector<uint32_t> p = lcg(7, 1000, true), q = lcg(8, 1000, true);
    std::vector<uint32_t> out(1000), want(1000);
    for (size_t i = 0; i < want.size(); ++i) want[i] = orc_bb_mul(p[i], q[i]);
    b.pointwise_mul(p, q, out);
    expect(out == want, "pointwise_mul BabyBear");
  }
  // errors: size checks before any GPU work, like the reference
  {
    Backend b = MetalCtGsR2(ctx);
    std::vector<uint32_t> bad(12);
    expect(throws<InvalidSize>([&] { b.forward_ntt(bad); }), "forward_ntt on 12 values throws InvalidSize");
    CosetLdeBatch lde(ctx);
    expect(throws<InvalidSize>([&] { lde.execute(std::vector<uint32_t>(100), 5, 3, 1); }),
           "CosetLdeBatch.execute on a short buffer throws InvalidSize");
    expect(throws<InvalidSize>([&] { lde.execute(std::vector<uint32_t>(size_t(1) << 26), 26, 1, 1); }),
           "LDE beyond two-adicity throws InvalidSize");
  }
  TwiddleCache tc(ctx);
  tc.prepare(Field::M31, 14);
  expect(tc.bytes() > 0, "TwiddleCache keeps tables resident");
  return failures ? 1 : 0;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc == 2 && std::string(argv[1]) == "--no-device") return no_device();
  if (argc == 2 && std::string(argv[1]) == "--parity") return parity();
  std::fprintf(stderr, "usage: %s --no-device | --parity\n", argv[0]);
  return 2;
}
