// Cube I/O and layouts.  ANA f0 files (uncompressed): the header block of
// ana_fzwrite / ana_fzread (cmlib/ana_write.cc:120-171, ana_read.cc:17-140:
// 512-byte blocks with the 0x5555aaaa synch pattern, subf bit 0 compressed
// and bit 7 big-endian data, nhb header blocks, data type, ndim and the
// dimensions with the first fastest).  The 12 x NX x NY x NZ model cube of
// atmosphere::read_spinor3d (atmos/atmos_fio.cc:79-127) to column states,
// and column Stokes spectra to the (nx, ny, 4, nlambda) output cube of the
// job (master/job.cc:590-597).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "rt_common.h"

namespace {

using snapi::kPi;

struct AnaHead {  // one 512-byte header block
  uint32_t synch;
  uint8_t subf, source, nhb, datyp, ndim, file_class;
  uint8_t cbytes[4];
  uint8_t free_[178];
  int32_t dim[16];
  char txt[256];
};
static_assert(sizeof(AnaHead) == 512, "ANA header block is 512 bytes");

const int kAnaSize[6] = {1, 2, 4, 4, 8, 8};  // INT8 INT16 INT32 FLOAT32 FLOAT64 INT64

uint32_t bswap32(uint32_t v) { return __builtin_bswap32(v); }

void swap_bytes(uint8_t *p, size_t n, int sz) {
  if (sz == 1) return;
  for (size_t i = 0; i < n; ++i, p += sz)
    for (int a = 0, b = sz - 1; a < b; ++a, --b) {
      const uint8_t t = p[a];
      p[a] = p[b];
      p[b] = t;
    }
}

// Reads the header blocks; data starts at nhb * 512.  Returns SNAPI_OK or an
// error; swap: the data's byte order differs from this (little-endian) host.
int ana_head(FILE *f, AnaHead &h, bool &swap, bool &compressed) {
  if (fread(&h, sizeof(h), 1, f) != 1) return SNAPI_ERR_ARG;
  const bool sync = h.synch == 0x5555aaaaU, rsync = h.synch == 0xaaaa5555U;
  if (!sync && !rsync) return SNAPI_ERR_ARG;
  if (h.nhb < 1 || h.nhb > 15 || h.ndim < 1 || h.ndim > 16 || h.datyp > 5) return SNAPI_ERR_UNSUPPORTED;
  if (rsync)  // a header written on a big-endian machine
    for (int d = 0; d < h.ndim; ++d) h.dim[d] = (int32_t)bswap32((uint32_t)h.dim[d]);
  const bool big = h.subf >= 128;  // data byte order (ana_read.cc:101-103)
  swap = big;
  if (rsync) swap = !swap;
  compressed = h.subf & 1;
  return SNAPI_OK;
}

// read_spinor3d's mapping of one model point (atmos_fio.cc:104-122) and the
// projection on the ray (atmosphere::project / transform, atmos_rad.cc:15-51).
template <typename R>
__global__ void cube_to_state_kernel(const double *__restrict__ cube, int nx, int ny, int nz, int nf, double theta,
                                     double phi, R *__restrict__ state) {
  const long npt = (long)nx * ny * nz;
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npt) return;
  const long col = p / nz, z = p % nz;  // col = x * ny + y
  auto q = [&](int k) { return cube[(long)k * npt + p]; };  // parameter k (0-based) of this point
  const double kB = 1.3806503E-16;
  const double T = q(2) > 3200.0 ? q(2) : 3200.0;
  const double B = q(7), el = q(10), az = q(11);
  const double Bx = B * sin(el) * cos(az), By = B * sin(el) * sin(az), Bz = B * cos(el);
  const double Vz = q(9);
  const double w1 = sin(phi) * sin(kPi - theta), w2 = cos(phi) * sin(kPi - theta), w3 = cos(kPi - theta);
  const double cp = cos(phi), sp = sin(phi), ct = cos(kPi - theta), st = sin(kPi - theta);
  const double mg = sqrt(Bx * Bx + By * By + Bz * Bz);
  const double nBx = ct * cp * Bx + sp * By - st * cp * Bz;
  const double nBy = -ct * sp * Bx + cp * By + st * sp * Bz;
  const double nBz = st * Bx + ct * Bz;
  R *s = state + col * nf * nz + z;
  for (int f = 0; f < nf; ++f) s[(long)f * nz] = R(0);
  s[SNAPI_F_GRID * nz] = (R)(-pow(10.0, q(0)));
  s[SNAPI_F_T * nz] = (R)T;
  s[SNAPI_F_NT * nz] = (R)(q(3) / (kB * T));
  s[SNAPI_F_NE * nz] = (R)(q(4) / (kB * T));
  s[SNAPI_F_OPREF * nz] = (R)q(6);
  s[SNAPI_F_VT * nz] = (R)q(8);
  s[SNAPI_F_VLOS * nz] = (R)(w1 * 0.0 + w2 * 0.0 + w3 * Vz);
  s[SNAPI_F_BMAG * nz] = (R)mg;
  s[SNAPI_F_BINC * nz] = (R)(acos(nBz / mg) * 180.0 / kPi);
  s[SNAPI_F_BAZI * nz] = (R)(atan(nBy / nBx) * 180.0 / kPi);
}

// stokes [ncol][nl][4] (col = x * ny + y) -> cube [nx][ny][4][nl], double.
template <typename R>
__global__ void stokes_to_cube_kernel(const R *__restrict__ stokes, long ncol, int nl, double *__restrict__ cube) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncol * nl * 4) return;
  const long col = i / (4L * nl), r = i % (4L * nl), s = r / nl, l = r % nl;
  cube[i] = (double)stokes[(col * nl + l) * 4 + s];
}

// Optical-depth scale of one column along a ray (atmos_ppbez::
// optical_depth_scale, atmos/plpar/atmos_ppbez_rts.cc:221-266): the Bezier
// integral of the opacity over the height grid with Fritsch-Butland
// derivatives (one-sided at the ends), tau = 1e-8 at the first point, the
// integration running downward for cos(theta) > 0 and upward otherwise.
// One work-item per column; the depth loop is serial, and each depth reads
// its neighbours' opacity and height once (HBM-bound, about 24 B per point).
__global__ void tau_scale_kernel(int ncol, int nd, const double *__restrict__ x3, const double *__restrict__ op,
                                 double ct, double sign, double *__restrict__ tau) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncol) return;
  const double *x = x3 + (size_t)col * nd, *o = op + (size_t)col * nd;
  double *t = tau + (size_t)col * nd;
  const int dz = ct > 0.0 ? 1 : -1, zl = ct > 0.0 ? 0 : nd - 1, zh = ct > 0.0 ? nd - 1 : 0;
  const double m = -ct;
  // Fritsch-Butland derivative of the opacity along the ray at point z
  auto deriv = [&](int z) {
    if (z == zl) return (o[zl + dz] - o[zl]) / (x[zl + dz] - x[zl]) * m;
    if (z == zh) return (o[zh] - o[zh - dz]) / (x[zh] - x[zh - dz]) * m;
    const double h2 = (x[z + dz] - x[z]) / m, h1 = (x[z] - x[z - dz]) / m;
    const double d2 = (o[z + dz] - o[z]) / h2, d1 = (o[z] - o[z - dz]) / h1;
    if (!(d1 * d2 > 0.0)) return 0.0;
    const double al = (1.0 + h1 / (h1 + h2)) / 3.0;
    return d1 * d2 / (al * d1 + (1.0 - al) * d2);
  };
  double tt = 1E-8, Dp = deriv(zl);
  t[zl] = sign * tt;
  for (int z = zl + dz; (z - zh - dz) * dz < 0; z += dz) {
    const double D = deriv(z);
    const double h = (x[z] - x[z - dz]) / m;
    const double C = 0.5 * ((o[z] - h / 2.0 * D) + (o[z - dz] + h / 2.0 * Dp));
    tt += h / 3.0 * (o[z - dz] + o[z] + C);
    t[z] = sign * tt;
    Dp = D;
  }
}

}  // namespace

extern "C" int snapi_optical_depth_scale_dev(int ncol, int nd, const double *height_dev, const double *op_dev,
                                             double theta, int referent, double *tau_dev, void *stream) {
  if (ncol <= 0 || nd < 2 || !height_dev || !op_dev || !tau_dev) return SNAPI_ERR_ARG;
  const double ct = cos(theta);
  if (ct == 0.0) return SNAPI_ERR_ARG;
  hipLaunchKernelGGL(tau_scale_kernel, dim3((ncol + 255) / 256), dim3(256), 0, static_cast<hipStream_t>(stream), ncol,
                     nd, height_dev, op_dev, ct, referent ? -1.0 : 1.0, tau_dev);
  return hipGetLastError() == hipSuccess ? SNAPI_OK : SNAPI_ERR_DEVICE;
}

extern "C" int snapi_optical_depth_scale(int ncol, int nd, const double *height, const double *op, double theta,
                                         int referent, double *tau) {
  if (ncol <= 0 || nd < 2 || !height || !op || !tau) return SNAPI_ERR_ARG;
  const size_t n = (size_t)ncol * nd * sizeof(double);
  double *dh = nullptr, *dop = nullptr, *dt = nullptr;
  int rc = SNAPI_OK;
  if (hipMalloc((void **)&dh, n) != hipSuccess || hipMalloc((void **)&dop, n) != hipSuccess ||
      hipMalloc((void **)&dt, n) != hipSuccess || hipMemcpy(dh, height, n, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dop, op, n, hipMemcpyHostToDevice) != hipSuccess)
    rc = SNAPI_ERR_DEVICE;
  if (rc == SNAPI_OK) rc = snapi_optical_depth_scale_dev(ncol, nd, dh, dop, theta, referent, dt, nullptr);
  if (rc == SNAPI_OK &&
      (hipDeviceSynchronize() != hipSuccess || hipMemcpy(tau, dt, n, hipMemcpyDeviceToHost) != hipSuccess))
    rc = SNAPI_ERR_DEVICE;
  hipFree(dh);
  hipFree(dop);
  hipFree(dt);
  return rc;
}

extern "C" int snapi_ana_info(const char *path, int *type, int *ndim, int *dims, int *compressed) {
  if (!path || !type || !ndim || !dims) return SNAPI_ERR_ARG;
  FILE *f = fopen(path, "rb");
  if (!f) return SNAPI_ERR_ARG;
  AnaHead h;
  bool swap, comp;
  const int rc = ana_head(f, h, swap, comp);
  fclose(f);
  if (rc) return rc;
  *type = h.datyp;
  *ndim = h.ndim;
  for (int d = 0; d < h.ndim; ++d) dims[d] = h.dim[d];
  if (compressed) *compressed = comp;
  return SNAPI_OK;
}

extern "C" int snapi_ana_read(const char *path, void *data, size_t bytes) {
  if (!path || !data) return SNAPI_ERR_ARG;
  FILE *f = fopen(path, "rb");
  if (!f) return SNAPI_ERR_ARG;
  AnaHead h;
  bool swap, comp;
  int rc = ana_head(f, h, swap, comp);
  if (!rc && comp) rc = SNAPI_ERR_UNSUPPORTED;  // crunch-compressed integer data is not read here
  size_t n = 1;
  for (int d = 0; !rc && d < h.ndim; ++d) n *= (size_t)h.dim[d];
  const int sz = rc ? 0 : kAnaSize[h.datyp];
  if (!rc && n * sz != bytes) rc = SNAPI_ERR_ARG;
  if (!rc && (fseek(f, (long)h.nhb * 512, SEEK_SET) != 0 || fread(data, 1, bytes, f) != bytes)) rc = SNAPI_ERR_ARG;
  fclose(f);
  if (!rc && swap) swap_bytes((uint8_t *)data, n, sz);
  return rc;
}

extern "C" int snapi_ana_write(const char *path, const void *data, int type, int ndim, const int *dims,
                               const char *header) {
  if (!path || !data || !dims || type < 0 || type > 5 || ndim < 1 || ndim > 16) return SNAPI_ERR_ARG;
  const size_t hl = header ? strlen(header) : 0;
  const int nhead = header ? (int)std::min<size_t>((hl + 254) / 255, 15) : 1;
  std::vector<AnaHead> h(nhead > 0 ? nhead : 1);
  memset(h.data(), 0, h.size() * sizeof(AnaHead));
  h[0].synch = 0x5555aaaaU;  // little-endian host
  h[0].nhb = (uint8_t)h.size();
  h[0].datyp = (uint8_t)type;
  h[0].ndim = (uint8_t)ndim;
  size_t n = 1;
  for (int d = 0; d < ndim; ++d) {
    if (dims[d] <= 0) return SNAPI_ERR_ARG;
    n *= (size_t)(h[0].dim[d] = dims[d]);
  }
  const size_t size = n * kAnaSize[type];
  const int res = (int)size;  // ana_fzwrite stores the data size as an int
  memcpy(h[0].cbytes, &res, 4);
  const char *p = header;
  for (size_t i = 0; header && i < h.size(); ++i) {
    const size_t len = std::min<size_t>(strlen(p), 255);
    memcpy(h[i].txt, p, len);
    h[i].txt[len] = 0;
    p += len;
  }
  FILE *f = fopen(path, "wb");
  if (!f) return SNAPI_ERR_ARG;
  int rc = SNAPI_OK;
  if (fwrite(h.data(), sizeof(AnaHead), h.size(), f) != h.size() || fwrite(data, 1, size, f) != size)
    rc = SNAPI_ERR_ARG;
  if (fclose(f) != 0) rc = SNAPI_ERR_ARG;
  return rc;
}

extern "C" int snapi_cube_to_state_dev(int precision, const double *cube_dev, int nx, int ny, int nz, int nfield,
                                       double theta, double phi, void *state_dev, void *stream) {
  if (!cube_dev || !state_dev || nx <= 0 || ny <= 0 || nz <= 0 || nfield < SNAPI_NF_BASE)
    return SNAPI_ERR_ARG;
  const long npt = (long)nx * ny * nz;
  const long blocks = (npt + 255) / 256;
  if (blocks > 0x7fffffffL) return SNAPI_ERR_UNSUPPORTED;
  hipStream_t st = (hipStream_t)stream;
  if (precision == SNAPI_FP64)
    hipLaunchKernelGGL(cube_to_state_kernel<double>, dim3((unsigned)blocks), dim3(256), 0, st, cube_dev, nx, ny, nz,
                       nfield, theta, phi, (double *)state_dev);
  else if (precision == SNAPI_MIXED)
    hipLaunchKernelGGL(cube_to_state_kernel<float>, dim3((unsigned)blocks), dim3(256), 0, st, cube_dev, nx, ny, nz,
                       nfield, theta, phi, (float *)state_dev);
  else
    return SNAPI_ERR_ARG;
  return hipGetLastError() == hipSuccess ? SNAPI_OK : SNAPI_ERR_DEVICE;
}

extern "C" int snapi_stokes_to_cube_dev(int precision, const void *stokes_dev, int nx, int ny, int nlambda,
                                        double *cube_dev, void *stream) {
  if (!stokes_dev || !cube_dev || nx <= 0 || ny <= 0 || nlambda <= 0) return SNAPI_ERR_ARG;
  const long ncol = (long)nx * ny, n = ncol * nlambda * 4;
  const long blocks = (n + 255) / 256;
  if (blocks > 0x7fffffffL) return SNAPI_ERR_UNSUPPORTED;
  hipStream_t st = (hipStream_t)stream;
  if (precision == SNAPI_FP64)
    hipLaunchKernelGGL(stokes_to_cube_kernel<double>, dim3((unsigned)blocks), dim3(256), 0, st,
                       (const double *)stokes_dev, ncol, nlambda, cube_dev);
  else if (precision == SNAPI_MIXED)
    hipLaunchKernelGGL(stokes_to_cube_kernel<float>, dim3((unsigned)blocks), dim3(256), 0, st,
                       (const float *)stokes_dev, ncol, nlambda, cube_dev);
  else
    return SNAPI_ERR_ARG;
  return hipGetLastError() == hipSuccess ? SNAPI_OK : SNAPI_ERR_DEVICE;
}
