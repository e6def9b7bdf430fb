// ct_common.hip -- error reporting, version, and the device exclusive scan used by the
// count -> scan -> write scheme of every variable-length query.
#include <cstdarg>
#include <cstring>

#include "ct_internal.h"
#include "ct_primitives.h"

namespace ct {

namespace {
thread_local char g_err[512] = "";
}

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int hip_fail(hipError_t e, const char* what, const char* file, int line) {
  set_error("HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e), what, file, line);
  return CT_ERR_HIP;
}

// ------------------------------------------------------------------ scan
// Three-kernel scan: per-tile sums (kScanTile elements per block), one block scans the
// tile sums, then every tile rescans with its carry.  int64 throughout.
constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;

__global__ void __launch_bounds__(kBlock) k_scan_tiles(const int64_t* __restrict__ in, int64_t n,
                                                       int64_t* __restrict__ tile_sums) {
  __shared__ int64_t lds[kBlock / 64];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int64_t s = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j)
    if (base + j < n) s += in[base + j];
  int64_t total;
  block_exclusive_scan<int64_t>(s, lds, &total);
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

__global__ void __launch_bounds__(kBlock) k_scan_carry(int64_t* __restrict__ tile_sums,
                                                       int64_t n_tiles) {
  __shared__ int64_t lds[kBlock / 64];
  int64_t carry = 0;
  for (int64_t base = 0; base < n_tiles; base += kBlock) {
    const int64_t i = base + threadIdx.x;
    const int64_t v = (i < n_tiles) ? tile_sums[i] : 0;
    int64_t total;
    const int64_t ex = block_exclusive_scan<int64_t>(v, lds, &total);
    if (i < n_tiles) tile_sums[i] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) tile_sums[n_tiles] = carry;
}

__global__ void __launch_bounds__(kBlock) k_scan_write(const int64_t* __restrict__ in, int64_t n,
                                                       const int64_t* __restrict__ tile_sums,
                                                       int64_t n_tiles,
                                                       int64_t* __restrict__ out) {
  __shared__ int64_t lds[kBlock / 64];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int64_t v[kScanItems];
  int64_t s = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    v[j] = (base + j < n) ? in[base + j] : 0;
    s += v[j];
  }
  int64_t total;
  int64_t run = tile_sums[blockIdx.x] + block_exclusive_scan<int64_t>(s, lds, &total);
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    if (base + j < n) out[base + j] = run;
    run += v[j];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = tile_sums[n_tiles];
}

int exclusive_scan(const int64_t* in, int64_t n, int64_t* out, void* ws, size_t ws_bytes,
                   hipStream_t st) {
  const int64_t n_tiles = (n + kScanTile - 1) / kScanTile;
  if (ws_bytes < ct_scan_workspace_bytes(n)) {
    set_error("scan workspace too small: %zu < %zu", ws_bytes, ct_scan_workspace_bytes(n));
    return CT_ERR_WORKSPACE;
  }
  int64_t* tiles = static_cast<int64_t*>(ws);
  if (n == 0) {
    return fill<int64_t>(out, 1, 0, st);
  }
  k_scan_tiles<<<(unsigned)n_tiles, kBlock, 0, st>>>(in, n, tiles);
  CT_LAUNCH_CHECK();
  k_scan_carry<<<1, kBlock, 0, st>>>(tiles, n_tiles);
  CT_LAUNCH_CHECK();
  k_scan_write<<<(unsigned)n_tiles, kBlock, 0, st>>>(in, n, tiles, n_tiles, out);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

}  // namespace ct

extern "C" {

const char* ct_last_error(void) { return ct::g_err; }

int ct_version(void) { return 1; }

size_t ct_scan_workspace_bytes(int64_t n) {
  const int64_t n_tiles = (n + ct::kScanTile - 1) / ct::kScanTile;
  return (size_t)(n_tiles + 1) * sizeof(int64_t) + 256;
}

int ct_exclusive_scan(const int64_t* in, int64_t n, int64_t* out, void* workspace,
                      size_t workspace_bytes, void* stream) {
  CT_REQUIRE(n >= 0, "n must be >= 0");
  CT_REQUIRE(out != nullptr && (n == 0 || in != nullptr), "null pointer");
  return ct::exclusive_scan(in, n, out, workspace, workspace_bytes, ct::as_stream(stream));
}

}  // extern "C"
