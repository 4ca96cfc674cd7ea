// Device-wide exclusive prefix sum of int32 values (reduce-then-scan):
// tile sums, one-workgroup scan of the tile sums, then a rescan of each tile
// with its offset.  out has n + 1 entries; out[n] is the total.
#include "scan.hpp"

#include "device_common.hpp"

namespace indigo {

namespace {

constexpr int kScanBlock = 512;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;

// Exclusive scan of one value per thread over the workgroup; returns the
// workgroup total through *total.
IG_DEV int block_exclusive_scan(int x, int* total)
{
  __shared__ int s_wave[kScanBlock / kWave];
  const int lane = lane_id();
  const int wid = threadIdx.x / kWave;
  const int incl = wave_inclusive_scan(x);
  if (lane == kWave - 1) s_wave[wid] = incl;
  __syncthreads();
  if (wid == 0) {
    constexpr int nw = kScanBlock / kWave;
    const int w = lane < nw ? s_wave[lane] : 0;
    const int wi = wave_inclusive_scan(w);
    if (lane < nw) s_wave[lane] = wi - w;
    if (lane == nw - 1) *total = wi;
  }
  __syncthreads();
  const int r = s_wave[wid] + incl - x;
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(kScanBlock) void tile_sums(const int* __restrict__ in, int n, int* __restrict__ sums)
{
  const long long base = static_cast<long long>(blockIdx.x) * kScanTile;
  int s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    const long long i = base + static_cast<long long>(k) * kScanBlock + threadIdx.x;
    if (i < n) s += in[i];
  }
  s = wave_sum(s);
  __shared__ int s_w[kScanBlock / kWave];
  if (lane_id() == 0) s_w[threadIdx.x / kWave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < kScanBlock / kWave; w++) t += s_w[w];
    sums[blockIdx.x] = t;
  }
}

// One workgroup scans the tile sums in place (exclusive); sums[m] = total.
__global__ __launch_bounds__(kScanBlock) void scan_sums(int* sums, int m)
{
  __shared__ int s_carry;
  __shared__ int s_tot;
  if (threadIdx.x == 0) s_carry = 0;
  __syncthreads();
  for (int base = 0; base < m; base += kScanBlock) {
    const int i = base + static_cast<int>(threadIdx.x);
    const int x = i < m ? sums[i] : 0;
    int tot = 0;
    const int ex = block_exclusive_scan(x, &s_tot);
    tot = s_tot;
    const int carry = s_carry;
    if (i < m) sums[i] = carry + ex;
    __syncthreads();
    if (threadIdx.x == 0) s_carry = carry + tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) sums[m] = s_carry;
}

__global__ __launch_bounds__(kScanBlock) void tile_scan(const int* __restrict__ in, int n, const int* __restrict__ sums,
                                                        int* __restrict__ out)
{
  __shared__ int s_tot;
  const long long base = static_cast<long long>(blockIdx.x) * kScanTile + static_cast<long long>(threadIdx.x) * kScanItems;
  int v[kScanItems];
  int s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    const long long i = base + k;
    v[k] = i < n ? in[i] : 0;
    s += v[k];
  }
  int run = block_exclusive_scan(s, &s_tot) + sums[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    const long long i = base + k;
    if (i < n) out[i] = run;
    run += v[k];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = sums[gridDim.x];
}

}  // namespace

size_t scan_temp_bytes(int n) { return static_cast<size_t>((n + kScanTile - 1) / kScanTile + 2) * sizeof(int); }

int exclusive_scan(hipStream_t stream, const int* in, int* out, int n, int* temp)
{
  if (n <= 0) return set_error(INDIGO_ERR_INVALID_ARG, "exclusive_scan: n must be positive");
  const int tiles = (n + kScanTile - 1) / kScanTile;
  tile_sums<<<tiles, kScanBlock, 0, stream>>>(in, n, temp);
  scan_sums<<<1, kScanBlock, 0, stream>>>(temp, tiles);
  tile_scan<<<tiles, kScanBlock, 0, stream>>>(in, n, temp, out);
  IG_HIP_TRY(hipGetLastError());
  return INDIGO_OK;
}

}  // namespace indigo
