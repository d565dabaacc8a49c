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

// Exclusive scan of n ints (single-pass, decoupled look-back); *total (host)
// receives the sum.  Integer sums are exact in any order, so the result is
// deterministic.
void scan_exclusive(Ctx& c, const int* in, int* out, int n, int* total) {
  if (n <= 0) {
    if (total) *total = 0;
    return;
  }
  size_t tmp = 0;
  CAC_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, in, out, n, c.stream));
  c.scan_tmp.reserve(tmp / sizeof(int) + 16);
  CAC_HIP(hipcub::DeviceScan::ExclusiveSum(c.scan_tmp.get(), tmp, in, out, n, c.stream));
  if (total) *total = read_device_sum(c, in + n - 1, out + n - 1);
}

// slot i opens the cells after the previous slot's cell up to its own
__global__ void k_cell_bounds(const int* __restrict__ key, int S, long ncells, int* __restrict__ cell_start) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > S) return;
  const long prev = i > 0 ? key[i - 1] : -1;
  const long cur = i < S ? key[i] : ncells;
  for (long c = prev + 1; c <= cur; c++) cell_start[c] = (int)i;
}

void sort_sites_by_cell(Ctx& c, const int* tcell, const unsigned long long* tval, int S) {
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
__device__ __forceinline__ unsigned long long spread3(unsigned v) {
  unsigned long long x = v & 0x1fffffull;
  x = (x | (x << 32)) & 0x1f00000000ffffull;
  x = (x | (x << 16)) & 0x1f0000ff0000ffull;
  x = (x | (x << 8)) & 0x100f00f00f00f00full;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}

__global__ void k_row_keys(const int* __restrict__ row_site, const int* __restrict__ scell, int R,
                           unsigned long long* __restrict__ key, int* __restrict__ idx) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  const int cc = scell[row_site[r]];
  key[r] = spread3((unsigned)cell_ix(cc)) | (spread3((unsigned)cell_iy(cc)) << 1) |
           (spread3((unsigned)cell_iz(cc)) << 2);
  idx[r] = r;
}

__global__ void k_row_gather(const int* __restrict__ perm, int R, const int* __restrict__ rs_in,
                             const int* __restrict__ rp_in, int* __restrict__ rs_out, int* __restrict__ rp_out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= R) return;
  const int r = perm[i];
  rs_out[i] = rs_in[r];
  rp_out[i] = rp_in[r];
}

void order_rows(Ctx& c) {
  const int R = c.P;
  c.row_key.reserve(2 * (size_t)R);
  c.row_idx.reserve(2 * (size_t)R);
  unsigned long long* key = c.row_key.get();
  int* idx = c.row_idx.get();
  k_row_keys<<<grid_for(R, 256), 256, 0, c.stream>>>(c.row_site.get(), c.scell.get(), R, key, idx);
  // the Morton key holds 3 bits per bit of the largest cell index
  int b = 1;
  while (b < 21 && (1L << b) < std::max(c.nc[0], std::max(c.nc[1], c.nc[2]))) b++;
  const int bits = 3 * b;
  // onesweep LSD passes over the key's bits at every size: below a million
  // items the library would otherwise merge-sort, in about 20 launches
  using OnesweepOnly = rocprim::radix_sort_config<rocprim::default_config, rocprim::default_config,
                                                  rocprim::default_config, 0>;
  size_t tmp = 0;
  CAC_HIP(rocprim::radix_sort_pairs<OnesweepOnly>(nullptr, tmp, key, key + R, idx, idx + R, (unsigned)R, 0u,
                                                  (unsigned)bits, c.stream));
  c.sort_tmp.reserve(tmp + 16);
  CAC_HIP(rocprim::radix_sort_pairs<OnesweepOnly>(c.sort_tmp.get(), tmp, key, key + R, idx, idx + R, (unsigned)R,
                                                  0u, (unsigned)bits, c.stream));
  c.row_tmp.reserve(2 * (size_t)R);
  int* rs = c.row_tmp.get();
  int* rp = rs + R;
  k_row_gather<<<grid_for(R, 256), 256, 0, c.stream>>>(idx + R, R, c.row_site.get(), c.row_point.get(), rs, rp);
  CAC_HIP(hipMemcpyAsync(c.row_site.get(), rs, sizeof(int) * R, hipMemcpyDeviceToDevice, c.stream));
  CAC_HIP(hipMemcpyAsync(c.row_point.get(), rp, sizeof(int) * R, hipMemcpyDeviceToDevice, c.stream));
}

}  // namespace cac
