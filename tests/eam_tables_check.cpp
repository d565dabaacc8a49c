// Driver for the CPU test of the EAM table builders (cac_amd/csrc/eam_tables.hpp):
// reads nr, dr, rhor[nr], z2r[nr] (int32, float64...) from argv[1], writes the
// splines of interpolate() and the mixed path's force table to argv[2]:
// rhor spline [(nr+1)*7] f64, z2r spline [(nr+1)*7] f64, table [3*(nr+2)*4] f32.
#include <cstdio>
#include <vector>

#include "eam_tables.hpp"

int main(int argc, char** argv) {
  if (argc != 3) return 2;
  FILE* in = std::fopen(argv[1], "rb");
  if (!in) return 3;
  int nr = 0;
  double dr = 0.0;
  if (std::fread(&nr, sizeof nr, 1, in) != 1 || std::fread(&dr, sizeof dr, 1, in) != 1 || nr < 5) return 4;
  std::vector<double> rho(nr + 1, 0.0), z2(nr + 1, 0.0);
  if (std::fread(rho.data() + 1, sizeof(double), nr, in) != (size_t)nr) return 5;
  if (std::fread(z2.data() + 1, sizeof(double), nr, in) != (size_t)nr) return 5;
  std::fclose(in);
  std::vector<double> rs, zs;
  cac::interpolate(nr, dr, rho, rs);
  cac::interpolate(nr, dr, z2, zs);
  std::vector<float> t;
  cac::mixed_force_table(nr, dr, rs, zs, t);
  FILE* out = std::fopen(argv[2], "wb");
  if (!out) return 6;
  std::fwrite(rs.data(), sizeof(double), rs.size(), out);
  std::fwrite(zs.data(), sizeof(double), zs.size(), out);
  std::fwrite(t.data(), sizeof(float), t.size(), out);
  return std::fclose(out) == 0 ? 0 : 7;
}
