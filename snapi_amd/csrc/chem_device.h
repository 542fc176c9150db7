// Device-side hydrostatic march and equation-of-state pass for callers that
// run them many times on batches of columns (chemeq.hip implements them).
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/snapi_gpu.h"

namespace snapi {

struct ChemDevice {
  void *buf = nullptr;  // table + derived constants
  void *map = nullptr;  // snapi_state_map, or null
  int nslot = 0;        // LDS doubles per team
  int nrec = 0;         // doubles per equilibrium record of the state pass
  int nf = 0;           // fields of the state the map fills
  alignas(8) unsigned char want[160];  // stage slots of the state pass
};

int chem_upload(const snapi_chem_table *tb, const snapi_state_map *map, ChemDevice &cd, hipStream_t st);
void chem_free(ChemDevice &cd);
// doubles of scratch chem_eos needs for npt depth points, in bytes
size_t chem_eos_scratch(const ChemDevice &cd, long npt, bool derivatives);
// hydrostatic equilibrium of ncol columns; with active, column c runs when
// active[c / group] is set
int chem_hequilibrium(const ChemDevice &cd, long ncol, int nd, const double *T, const double *logtau, double *Nt,
                      const int *active, int group, hipStream_t st);
// tangent of the hydrostatic equilibrium of ncol converged columns along
// ndir temperature perturbations each (snapi_hequilibrium_tangent): dT, dN
// [ncol][ndir][nd]; scratch of chem_he_tangent_scratch bytes; columns with
// active[c] == 0 are skipped
size_t chem_he_tangent_scratch(long ncol, int nd);
int chem_he_tangent(const ChemDevice &cd, long ncol, int nd, const double *T, const double *logtau, const double *Nt,
                    int ndir, const double *dT, double *dN, double *scratch, const int *active, hipStream_t st);
// the equation-of-state pass (snapi_eos_state) on ncol columns of a double
// state; resp/dstate may be null; columns with active[c] == 0 are skipped
int chem_eos(const ChemDevice &cd, long ncol, int nd, double *state, double *resp, double *dstate, double *scratch,
             const int *active, hipStream_t st);
// the same pass with derivatives, whose first five chemical equilibria per
// point also give the hydrostatic tangent's partials at cd_he's referent
// wavelength (he_scratch: chem_he_tangent_scratch bytes), then that tangent
// (as chem_he_tangent) at the state's T and Nt, given as T, Nt [ncol][nd]
int chem_eos_he_tangent(const ChemDevice &cd, const ChemDevice &cd_he, long ncol, int nd, double *state,
                        double *resp, double *dstate, double *scratch, const double *T, const double *logtau,
                        const double *Nt, int ndir, const double *dT, double *dN, double *he_scratch,
                        const int *active, hipStream_t st);

// Gaussian spectral convolution of nseries [nl][ncomp] series in place
// (instrument.hip); lam_h the host copy of lam_d; series with active == 0
// are skipped.
int spectral_convolve_launch(int precision, int nseries, int nl, int ncomp, const double *lam_h,
                             const double *lam_d, double width, void *data, const int *active, hipStream_t st);
// Sampled-PSF convolution of the same series (instrument.hip); psf_d the
// npsf samples on the device.
int psf_convolve_launch(int precision, int nseries, int nl, int ncomp, int npsf, const double *psf_d, void *data,
                        const int *active, hipStream_t st);

}  // namespace snapi
