// sort.hip -- the cell ordering of the sites at a neighbor build.
//
// Replaces the binning of NBinStandard::bin_atoms (nbin_standard.cpp): every
// candidate site gets its cell, and the sites are laid out cell by cell.  A
// stable LSD radix sort of (cell, candidate index) pairs gives each cell's
// sites in candidate order, so the layout is deterministic; cell_start[c] is
// the first slot of cell c (cell_start[ncells] = S).
#include <hipcub/hipcub.hpp>

#include "cac_ctx.hpp"

namespace cac {

__global__ void k_iota(int* __restrict__ v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = i;
}

// slot i opens the cells after the previous slot's cell up to its own
__global__ void k_cell_bounds(const int* __restrict__ key, int S, long ncells, int* __restrict__ cell_start) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > S) return;
  const long prev = i > 0 ? key[i - 1] : -1;
  const long cur = i < S ? key[i] : ncells;
  for (long c = prev + 1; c <= cur; c++) cell_start[c] = (int)i;
}

void sort_sites_by_cell(Ctx& c, const int* tcell, int S, int* perm) {
  int bits = 1;
  while (bits < 31 && (1L << bits) < c.ncells) bits++;
  c.sort_keys.reserve(S);
  c.sort_vals.reserve(S);
  k_iota<<<grid_for(S, 256), 256, 0, c.stream>>>(c.sort_vals.get(), S);
  size_t tmp = 0;
  CAC_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, tcell, c.sort_keys.get(), c.sort_vals.get(), perm, S, 0,
                                             bits, c.stream));
  c.sort_tmp.reserve(tmp + 16);
  CAC_HIP(hipcub::DeviceRadixSort::SortPairs(c.sort_tmp.get(), tmp, tcell, c.sort_keys.get(), c.sort_vals.get(), perm,
                                             S, 0, bits, c.stream));
  c.cell_start.reserve(c.ncells + 1);
  k_cell_bounds<<<grid_for((long)S + 1, 256), 256, 0, c.stream>>>(c.sort_keys.get(), S, c.ncells,
                                                                  c.cell_start.get());
}

}  // namespace cac
