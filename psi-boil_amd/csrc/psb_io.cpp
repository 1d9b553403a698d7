// psb_io.cpp -- the reference's .bck checkpoint of a cell field
// (Scalar::save / Scalar::load, Src/Field/Scalar/scalar_save.cpp:20-32,
// scalar_load.cpp:29-39): three ints, the padded extents ni nj nk (interior
// plus bw ghost layers per side), then ni*nj*nk doubles of the padded array
// in [i][j][k] order, k fastest, native byte order.  The file name follows
// name_file (Src/Global/global_name_file.cpp): <name>_p<rank, 3 digits>_
// <step, 6 digits>.bck.  Host code only: the device variants in
// psb_api.hip copy through these.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/psb.h"

namespace {

size_t padded(int n, int bw) { return (size_t)n + 2 * (size_t)bw; }

}  // namespace

extern "C" {

int psb_bck_name(const char *name, int step, int rank, char *out, int cap) {
  if (!name || !out || cap < 1) return PSB_EINVAL;
  std::string s = name;
  char buf[32];
  if (rank > -1) {
    std::snprintf(buf, sizeof(buf), "_p%03d", rank);
    s += buf;
  }
  if (step > -1) {
    std::snprintf(buf, sizeof(buf), "_%06d", step);
    s += buf;
  }
  s += ".bck";
  if ((int)s.size() + 1 > cap) return PSB_EINVAL;
  std::memcpy(out, s.c_str(), s.size() + 1);
  return PSB_OK;
}

int psb_save_bck(const char *path, int nx, int ny, int nz, int bw, const double *x) {
  if (!path || !x || nx < 1 || ny < 1 || nz < 1 || bw < 0) return PSB_EINVAL;
  const int ni = (int)padded(nx, bw), nj = (int)padded(ny, bw), nk = (int)padded(nz, bw);
  // the ghost layers take the nearest interior value (a zero-gradient
  // update); a reader that refreshes its boundaries after loading, as the
  // reference does, uses the interior only
  std::vector<double> v((size_t)ni * nj * nk);
  auto clampi = [](int q, int n) { return q < 0 ? 0 : (q >= n ? n - 1 : q); };
  size_t o = 0;
  for (int i = 0; i < ni; i++)
    for (int j = 0; j < nj; j++)
      for (int k = 0; k < nk; k++) {
        const int a = clampi(i - bw, nx), b = clampi(j - bw, ny), c = clampi(k - bw, nz);
        v[o++] = x[((size_t)a * ny + b) * nz + c];
      }
  std::ofstream out(path, std::ios::binary);
  if (!out) return PSB_EINVAL;
  out.write(reinterpret_cast<const char *>(&ni), sizeof(int));
  out.write(reinterpret_cast<const char *>(&nj), sizeof(int));
  out.write(reinterpret_cast<const char *>(&nk), sizeof(int));
  out.write(reinterpret_cast<const char *>(v.data()), (std::streamsize)(v.size() * sizeof(double)));
  return out.good() ? PSB_OK : PSB_EINVAL;
}

int psb_load_bck(const char *path, int nx, int ny, int nz, double *x, int *bw_out) {
  if (!path || !x || nx < 1 || ny < 1 || nz < 1) return PSB_EINVAL;
  std::ifstream in(path, std::ios::binary);
  if (!in) return PSB_EINVAL;
  int n[3] = {0, 0, 0};
  in.read(reinterpret_cast<char *>(n), sizeof(n));
  if (!in) return PSB_EINVAL;
  // one ghost width on every side: the extents must exceed the interior
  // by the same even amount
  const int bw = (n[0] - nx) / 2;
  if (bw < 0 || n[0] != nx + 2 * bw || n[1] != ny + 2 * bw || n[2] != nz + 2 * bw)
    return PSB_EINVAL;
  std::vector<double> v((size_t)n[0] * n[1] * n[2]);
  in.read(reinterpret_cast<char *>(v.data()), (std::streamsize)(v.size() * sizeof(double)));
  if (!in) return PSB_EINVAL;
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++)
        x[((size_t)i * ny + j) * nz + k] =
            v[((size_t)(i + bw) * n[1] + (j + bw)) * n[2] + (k + bw)];
  if (bw_out) *bw_out = bw;
  return PSB_OK;
}

}  // extern "C"
