// sort.hip -- the cell ordering of the sites at a neighbor build.
//
// Replaces the binning of NBinStandard::bin_atoms (nbin_standard.cpp): every
// candidate site gets its cell, and the sites are laid out cell by cell.  A
// stable LSD radix sort of (cell, packed site descriptor) pairs gives each
// cell's sites in candidate order, so the layout is deterministic, and moves
// the descriptors with the keys (no gather through a permutation afterwards);
// cell_start[c] is the first slot of cell c (cell_start[ncells] = S).
#include <algorithm>

#include <hipcub/hipcub.hpp>
#include <rocprim/device/device_radix_sort.hpp>

#include "cac_ctx.hpp"

namespace cac {

// ------------------------------------------------------------------ scans
// Exclusive scan of n small non-negative ints in three launches over tiles of
// SCAN_TILE elements: the tiles' sums, their exclusive scan in one workgroup
// (which also leaves the total on the device), then each tile's scan plus its
// offset.  Integer sums are exact in any order, so the result is
// deterministic.  The live-site variant also compacts the flagged sites in
// the last launch, which then reads each tile's flags once for both.
constexpr int SCAN_NT = 256;
constexpr int SCAN_IPT = 16;  // consecutive elements per thread
constexpr int SCAN_TILE = SCAN_NT * SCAN_IPT;

__device__ __forceinline__ int wave_incl_scan(int x) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  return x;
}

// exclusive prefix of x over the workgroup's NT threads; tot = their sum
template <int NT>
__device__ __forceinline__ int block_excl_scan(int x, int* lds, int& tot) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int inc = wave_incl_scan(x);
  if (lane == 63) lds[w] = inc;
  __syncthreads();
  int before = 0;
  tot = 0;
  for (int k = 0; k < NT / 64; k++) {
    const int s = lds[k];
    before += k < w ? s : 0;
    tot += s;
  }
  __syncthreads();  // lds is written again by the next call
  return before + inc - x;
}

// the flags of 16 sites as one 16-byte load
__device__ __forceinline__ void load_run(const unsigned char* __restrict__ in, long base, long n, int* v) {
  if (base + SCAN_IPT <= n) {
    const uint4 t = *reinterpret_cast<const uint4*>(in + base);
    const unsigned w[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
    for (int j = 0; j < SCAN_IPT; j++) v[j] = (int)((w[j >> 2] >> (8 * (j & 3))) & 0xffu);
  } else {
#pragma unroll
    for (int j = 0; j < SCAN_IPT; j++) v[j] = base + j < n ? (int)in[base + j] : 0;
  }
}

__device__ __forceinline__ void load_run(const int* __restrict__ in, long base, long n, int* v) {
  if (base + SCAN_IPT <= n) {
    const int4* p = reinterpret_cast<const int4*>(in + base);
#pragma unroll
    for (int q = 0; q < SCAN_IPT / 4; q++) {
      const int4 t = p[q];
      v[4 * q] = t.x;
      v[4 * q + 1] = t.y;
      v[4 * q + 2] = t.z;
      v[4 * q + 3] = t.w;
    }
  } else {
#pragma unroll
    for (int j = 0; j < SCAN_IPT; j++) v[j] = base + j < n ? in[base + j] : 0;
  }
}

template <typename T>
__global__ __launch_bounds__(SCAN_NT) void k_scan_tiles(const T* __restrict__ in, int n, int* __restrict__ tsum) {
  __shared__ int lds[SCAN_NT / 64];
  int v[SCAN_IPT];
  load_run(in, (long)blockIdx.x * SCAN_TILE + (long)threadIdx.x * SCAN_IPT, n, v);
  int s = 0;
#pragma unroll
  for (int j = 0; j < SCAN_IPT; j++) s += v[j];
  int tot;
  block_excl_scan<SCAN_NT>(s, lds, tot);
  if (threadIdx.x == 0) tsum[blockIdx.x] = tot;
}

// one workgroup: the tile sums become exclusive tile offsets in place
__global__ __launch_bounds__(1024) void k_scan_offsets(int* __restrict__ tsum, int nt, int* __restrict__ total) {
  __shared__ int lds[1024 / 64];
  int carry = 0;
  for (int b = 0; b < nt; b += 1024) {
    const int i = b + (int)threadIdx.x;
    const int x = i < nt ? tsum[i] : 0;
    int tot;
    const int ex = block_excl_scan<1024>(x, lds, tot);
    if (i < nt) tsum[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0 && total) *total = carry;
}

// MODE 0: plain scan; 1 / 2: live sites with FP32 / FP64 records
template <int MODE, typename T>
__global__ __launch_bounds__(SCAN_NT) void k_scan_apply(const T* __restrict__ in, int n,
                                                        const int* __restrict__ toff, int* __restrict__ out,
                                                        LiveOut lo) {
  __shared__ int lds[SCAN_NT / 64];
  __shared__ int sidx[MODE ? SCAN_TILE : 1];
  const long tile0 = (long)blockIdx.x * SCAN_TILE;
  const long base = tile0 + (long)threadIdx.x * SCAN_IPT;
  int v[SCAN_IPT];
  load_run(in, base, n, v);
  int s = 0;
#pragma unroll
  for (int j = 0; j < SCAN_IPT; j++) s += v[j];
  int tot;
  int run = toff[blockIdx.x] + block_excl_scan<SCAN_NT>(s, lds, tot);
  int o[SCAN_IPT];
#pragma unroll
  for (int j = 0; j < SCAN_IPT; j++) {
    o[j] = run;
    run += v[j];
  }
  if (base + SCAN_IPT <= n) {
    int4* p = reinterpret_cast<int4*>(out + base);
#pragma unroll
    for (int q = 0; q < SCAN_IPT / 4; q++) p[q] = make_int4(o[4 * q], o[4 * q + 1], o[4 * q + 2], o[4 * q + 3]);
  } else {
#pragma unroll
    for (int j = 0; j < SCAN_IPT; j++)
      if (base + j < n) out[base + j] = o[j];
  }
  if (MODE) {
    // the flags this thread read are cleared for the next prune
    if (base + SCAN_IPT <= n)
      *reinterpret_cast<uint4*>(lo.clear + base) = make_uint4(0u, 0u, 0u, 0u);
    else
      for (int j = 0; j < SCAN_IPT; j++)
        if (base + j < n) lo.clear[base + j] = 0;
    // each flagged site's dense index (-1 when unflagged), staged so that the
    // copies below run over consecutive sites across the workgroup
#pragma unroll
    for (int j = 0; j < SCAN_IPT; j++) sidx[threadIdx.x * SCAN_IPT + j] = v[j] ? o[j] : -1;
    __syncthreads();
    // batches of CB sites per thread: all loads, then all stores (the output
    // arrays could alias the inputs as far as the compiler knows, so loads
    // after a store would wait for it)
    constexpr int CB = 8;
    for (int j0 = 0; j0 < SCAN_IPT; j0 += CB) {
      int ii[CB], dd[CB], mm[CB], cc[CB];
      float4 r32[CB];
      double4 r64[CB];
#pragma unroll
      for (int u = 0; u < CB; u++) {
        const int e = (j0 + u) * SCAN_NT + (int)threadIdx.x;
        const long k = tile0 + e;
        ii[u] = k < n ? sidx[e] : -1;
        if (ii[u] >= 0) {
          dd[u] = lo.sdesc[k];
          mm[u] = lo.simg[k];
          cc[u] = lo.scell[k];
          if (MODE == 1)
            r32[u] = lo.site32f[k];
          else
            r64[u] = lo.site[k];
        }
      }
#pragma unroll
      for (int u = 0; u < CB; u++) {
        const int i = ii[u];
        if (i < 0) continue;
        lo.lsdesc[i] = dd[u];
        lo.lsimg[i] = mm[u];
        lo.lscell[i] = cc[u];
        if (MODE == 1)
          lo.site32[i] = make_float4(r32[u].x, r32[u].y, r32[u].z, 0.0f);
        else
          lo.site_live[i] = make_double4(r64[u].x, r64[u].y, r64[u].z, 0.0);
      }
    }
  }
}

template <typename T>
static void scan_launch(Ctx& c, const T* in, int* out, int n, int* d_total, const LiveOut* lo) {
  const int nt = (n + SCAN_TILE - 1) / SCAN_TILE;
  c.scan_tmp.reserve((size_t)nt + 1);
  if (nt > 0) k_scan_tiles<T><<<nt, SCAN_NT, 0, c.stream>>>(in, n, c.scan_tmp.get());
  k_scan_offsets<<<1, 1024, 0, c.stream>>>(c.scan_tmp.get(), nt, d_total);
  if (nt > 0) {
    if (!lo)
      k_scan_apply<0, T><<<nt, SCAN_NT, 0, c.stream>>>(in, n, c.scan_tmp.get(), out, LiveOut{});
    else if (lo->mixed)
      k_scan_apply<1, T><<<nt, SCAN_NT, 0, c.stream>>>(in, n, c.scan_tmp.get(), out, *lo);
    else
      k_scan_apply<2, T><<<nt, SCAN_NT, 0, c.stream>>>(in, n, c.scan_tmp.get(), out, *lo);
  }
  CAC_HIP(hipGetLastError());
}

// *total (host) receives the sum
void scan_exclusive(Ctx& c, const int* in, int* out, int n, int* total) {
  if (n <= 0) {
    if (total) *total = 0;
    return;
  }
  scan_launch(c, in, out, n, nullptr, nullptr);
  if (total) *total = read_device_sum(c, in + n - 1, out + n - 1);
}

void scan_live(Ctx& c, const unsigned char* flag, int* idx, int n, int* d_total, const LiveOut& lo) {
  scan_launch(c, flag, idx, n > 0 ? n : 0, d_total, &lo);
}

// slot i opens the cells after the previous slot's cell up to its own
__global__ void k_cell_bounds(const int* __restrict__ key, int S, long ncells, int* __restrict__ cell_start) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > S) return;
  const long prev = i > 0 ? key[i - 1] : -1;
  const long cur = i < S ? key[i] : ncells;
  for (long c = prev + 1; c <= cur; c++) cell_start[c] = (int)i;
}

void sort_sites_by_cell(Ctx& c, const int* tcell, const unsigned* tval, int S) {
  int bits = 1;
  while (bits < 31 && (1L << bits) < c.ncells) bits++;
  c.sort_keys.reserve(S);
  c.sort_vals.reserve(S);
  size_t tmp = 0;
  CAC_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, tcell, c.sort_keys.get(), tval, c.sort_vals.get(), S, 0,
                                             bits, c.stream));
  c.sort_tmp.reserve(tmp + 16);
  CAC_HIP(hipcub::DeviceRadixSort::SortPairs(c.sort_tmp.get(), tmp, tcell, c.sort_keys.get(), tval,
                                             c.sort_vals.get(), S, 0, bits, c.stream));
  c.cell_start.reserve(c.ncells + 1);
  k_cell_bounds<<<grid_for((long)S + 1, 256), 256, 0, c.stream>>>(c.sort_keys.get(), S, c.ncells,
                                                                  c.cell_start.get());
}

// Rows (points) in Morton order of their cells: consecutive rows, and so the
// rows one XCD works on at a time, form compact 3D blocks, whose neighbor sites
// fit the XCD's L2 far better than the slabs of the plain cell order.
//
// The rows arrive in site order, so each cell's rows are one contiguous run in
// cell order.  Reordering the runs by the Morton order of their cells (a
// permutation of the cells that depends only on the grid shape, sorted once
// per shape) gives the rows a stable sort by Morton key would give: a scan of
// the runs' lengths in that order, then one copy per run.
__device__ __forceinline__ unsigned long long spread3(unsigned v) {
  unsigned long long x = v & 0x1fffffull;
  x = (x | (x << 32)) & 0x1f00000000ffffull;
  x = (x | (x << 16)) & 0x1f0000ff0000ffull;
  x = (x | (x << 8)) & 0x100f00f00f00f00full;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}

__global__ void k_cell_morton_keys(int ncells, int nc0, int nc1, unsigned long long* __restrict__ key,
                                   int* __restrict__ idx) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  const unsigned ix = (unsigned)(c % nc0), iy = (unsigned)((c / nc0) % nc1), iz = (unsigned)(c / (nc0 * nc1));
  key[c] = spread3(ix) | (spread3(iy) << 1) | (spread3(iz) << 2);
  idx[c] = c;
}

// the run of cell c: rows [first, first + count), clamped to the P rows
__device__ __forceinline__ void row_run(const int* __restrict__ cell_start, const int* __restrict__ rowidx, int c,
                                        int S, int P, int& first, int& count) {
  const int a = cell_start[c], b = cell_start[c + 1];
  const int ra = min(a < S ? rowidx[a] : P, P), rb = min(b < S ? rowidx[b] : P, P);
  first = ra;
  count = max(rb - ra, 0);
}

__global__ void k_row_runs(const int* __restrict__ morton, const int* __restrict__ cell_start,
                           const int* __restrict__ rowidx, int ncells, int S, int P, int* __restrict__ count) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= ncells) return;
  int first, n;
  row_run(cell_start, rowidx, morton[m], S, P, first, n);
  count[m] = n;
}

__global__ void k_row_copy(const int* __restrict__ morton, const int* __restrict__ cell_start,
                           const int* __restrict__ rowidx, const int* __restrict__ start, int ncells, int S, int P,
                           const int* __restrict__ rs_in, const int* __restrict__ rp_in, int* __restrict__ rs_out,
                           int* __restrict__ rp_out) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= ncells) return;
  int first, n;
  row_run(cell_start, rowidx, morton[m], S, P, first, n);
  const int o = start[m];
  for (int j = 0; j < n && o + j < P; j++) {
    rs_out[o + j] = rs_in[first + j];
    rp_out[o + j] = rp_in[first + j];
  }
}

static void sort_cells_morton(Ctx& c) {
  const int nc = (int)c.ncells;
  if (c.morton_nc[0] == c.nc[0] && c.morton_nc[1] == c.nc[1] && c.morton_nc[2] == c.nc[2]) return;
  c.row_key.reserve(2 * (size_t)nc);
  c.row_idx.reserve(2 * (size_t)nc);
  unsigned long long* key = c.row_key.get();
  int* idx = c.row_idx.get();
  k_cell_morton_keys<<<grid_for(nc, 256), 256, 0, c.stream>>>(nc, c.nc[0], c.nc[1], key, idx);
  // the Morton key holds 3 bits per bit of the largest cell index
  int b = 1;
  while (b < 21 && (1L << b) < std::max(c.nc[0], std::max(c.nc[1], c.nc[2]))) b++;
  const int bits = 3 * b;
  size_t tmp = 0;
  CAC_HIP(rocprim::radix_sort_pairs(nullptr, tmp, key, key + nc, idx, idx + nc, (unsigned)nc, 0u, (unsigned)bits,
                                    c.stream));
  c.sort_tmp.reserve(tmp + 16);
  CAC_HIP(rocprim::radix_sort_pairs(c.sort_tmp.get(), tmp, key, key + nc, idx, idx + nc, (unsigned)nc, 0u,
                                    (unsigned)bits, c.stream));
  c.cell_morton.reserve(nc);
  CAC_HIP(hipMemcpyAsync(c.cell_morton.get(), idx + nc, sizeof(int) * nc, hipMemcpyDeviceToDevice, c.stream));
  for (int d = 0; d < 3; d++) c.morton_nc[d] = c.nc[d];
}

void order_rows(Ctx& c) {
  const int R = c.P, nc = (int)c.ncells;
  sort_cells_morton(c);
  c.morton_rows.reserve(2 * (size_t)nc);
  int* count = c.morton_rows.get();
  int* start = count + nc;
  k_row_runs<<<grid_for(nc, 256), 256, 0, c.stream>>>(c.cell_morton.get(), c.cell_start.get(), c.rowidx.get(), nc,
                                                      c.S, R, count);
  scan_launch(c, count, start, nc, nullptr, nullptr);
  c.row_tmp.reserve(2 * (size_t)R);
  int* rs = c.row_tmp.get();
  int* rp = rs + R;
  k_row_copy<<<grid_for(nc, 256), 256, 0, c.stream>>>(c.cell_morton.get(), c.cell_start.get(), c.rowidx.get(),
                                                      start, nc, c.S, R, c.row_site.get(), c.row_point.get(), rs,
                                                      rp);
  CAC_HIP(hipMemcpyAsync(c.row_site.get(), rs, sizeof(int) * R, hipMemcpyDeviceToDevice, c.stream));
  CAC_HIP(hipMemcpyAsync(c.row_point.get(), rp, sizeof(int) * R, hipMemcpyDeviceToDevice, c.stream));
}

}  // namespace cac
