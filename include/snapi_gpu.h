/*
 * snapi_gpu.h -- C ABI of the MI355X (gfx950) implementation of SNAPI's
 * data-parallel hot path: Zeeman opacity build, polarized formal solution,
 * per-depth responses (tangent) and their adjoint, the Levenberg-Marquardt
 * normal-equation step, and the scalar formal solution with approximate
 * Lambda operator used by the ALI iteration.
 *
 * Every entry point takes plain host pointers and sizes, copies inputs to the
 * device, runs the HIP kernels and copies results back (the *_dev variants
 * take device pointers and a hipStream_t given as void* instead and do no
 * copies).  Return value: SNAPI_OK (0) or a negative snapi_status.
 *
 * Reference interfaces replaced (ivanzmilic/snapi):
 *   snapi_synth_stokes*      atmosphere::obs_stokes                 atmos/atmos_obs.cc:16-87
 *   snapi_line_add, snapi_zeeman_pattern, snapi_vdw_damping, snapi_radiative_damping
 *                            the atom class's line data: atom::atom, damp_rad,
 *                            compute_damp_col, zeeman_setup  atmos/atmol/atom/atom.cc:233-253,
 *                                                                   926-935, 987-1308, 1620-1773
 *                            (atmosphere::op_em_vector              atmos/atmos_rad_synthesis.cc:20-72
 *                             + atmos_ppbez::formal (polarized)     atmos/plpar/atmos_ppbez_rts.cc:918-1218)
 *   snapi_stokes_responses*  atmosphere::obs_stokes_responses       atmos/atmos_rad_synthesis.cc:166-399
 *                            (op_em_vector_plus_pert + formal_pert_numerical,
 *                                                                   atmos/plpar/atmos_ppbez_rts.cc:1780-1857)
 *   snapi_stokes_adjoint     (no counterpart: gradient J^T r of the chi^2 of
 *                            atmos/atmos_fit.cc:229-248 in one reverse sweep)
 *   snapi_lm_step            the damped normal-equation solve of
 *                            atmosphere::stokes_lm_fit              atmos/atmos_fit.cc:129-160
 *                            (multiply_with_transpose, solve_svd    cmlib/mathtools.cc:615-624, 344-387)
 *   snapi_formal_scalar      atmos_ppbez::formal (scalar, with Lambda*)
 *                                                                   atmos/plpar/atmos_ppbez_rts.cc:21-219
 *   snapi_node_jacobian[_eos] atmosphere::obs_stokes_responses_to_nodes
 *                                                                   atmos/atmos_fit.cc:1205-1352
 *   snapi_invert*[_eos]      atmosphere::stokes_lm_fit              atmos/atmos_fit.cc:18-227
 *   snapi_build_from_nodes*  atmos_pp::build_from_nodes + popsetup  atmos/plpar/atmos_pp.cc:58-180
 *   snapi_ana_*              ana_fzread / ana_fzwrite               cmlib/ana_read.cc, ana_write.cc
 *   snapi_cube_to_state_dev  atmosphere::read_spinor3d (+ project)  atmos/atmos_fio.cc:79-127
 *   snapi_optical_depth_scale* atmos_ppbez::optical_depth_scale, compute_tau_referent
 *                                                                   atmos/plpar/atmos_ppbez_rts.cc:221-266, 328-338
 *   snapi_stokes_to_cube_dev the job's spectra cube                 master/job.cc:590-597
 *   snapi_spectral_convolve* convolve_spectra_with_gauss, convolve_response_with_gauss
 *                                                                   cmlib/mathtools.cc:1350-1427, 1473-1498
 *   snapi_psf_convolve*      convolve_spectra_with_psf, convolve_response_with_psf
 *                                                                   cmlib/mathtools.cc:1429-1465, 1500-1522
 *   snapi_add_scattered_light* observable::add_scattered_light       cmlib/obs.cc:412-428
 *   snapi_ali_accumulate     atom::add (Jbar, Lambda*, norm sums)   atmos/atmol/atom/atom.cc:1865-1927
 *   snapi_ali_newpops        atom::newpops                          atmos/atmol/atom/atom_pop.cc:27-135
 *   snapi_ali_solve*         atmosphere::nltepops (one NLTE atom)   atmos/atmos_pop.cc:67-240
 *   snapi_formal_responses*  atmos_ppbez::formal_with_responses    atmos/plpar/atmos_ppbez_rts.cc:349-778
 *   snapi_response_solve     the Crout solves of atom::compute_nlte_population_responses
 *                                                                   atmos/atmol/atom/atom_responses.cc:171-200
 *   snapi_nlte_response_system* atmosphere::compute_nlte_population_responses with
 *                            atom::compute_profile_norm, add_response_contributions_new and
 *                            the rate part of atom::compute_nlte_population_responses
 *                                                                   atmos/atmos_responses.cc:13-157,
 *                                                                   atmos/atmol/atom/atom_rad.cc:885-930,
 *                                                                   atmos/atmol/atom/atom_responses.cc:26-200, 668-1128
 *   snapi_chemeq*            atmosphere::ltepops: chemeq + atom::lte atmos/atmos_pop.cc:30-45,
 *                                                                   atmos/atmos_chm.cc:28-363,
 *                                                                   atmos/atmol/atom/atom.cc:779-787, 892-924
 *   snapi_hequilibrium*      atmos_pp::enforce_hequilibrium         atmos/plpar/atmos_pp.cc:372-455
 *   snapi_eos_state*         popsetup + compute_op_referent, and the response path's
 *                            compute_op_referent_derivative, ne_lte_derivatives and
 *                            compute_lte_population_responses_analytical
 *                                                                   atmos/atmos_rad_synthesis.cc:173-190
 */
#ifndef SNAPI_GPU_H
#define SNAPI_GPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum snapi_status {
  SNAPI_OK = 0,
  SNAPI_ERR_ARG = -1,         /* bad sizes or null pointers */
  SNAPI_ERR_DEVICE = -2,      /* HIP runtime error */
  SNAPI_ERR_UNSUPPORTED = -3, /* input outside what the kernels implement */
  SNAPI_ERR_NONFINITE = -4,   /* a result was NaN/Inf */
  SNAPI_ERR_CAPACITY = -5     /* the problem needs more than a fixed-size work area holds */
};

/* Precision of the compute path.
 * SNAPI_FP64: every operation in double, the reference's arithmetic.
 * SNAPI_MIXED: opacity build and formal solution in float (Taylor weights
 *              for optical-depth steps below 1e-2); the wavelength-independent
 *              per-depth setup, the line-centre offsets and the LM damping
 *              solve in double. */
enum snapi_precision { SNAPI_FP64 = 0, SNAPI_MIXED = 1 };

/*
 * Column state ("column tile"): the per-depth quantities the reference's
 * atmosphere object holds after popsetup()/compute_op_referent()/nltepops(),
 * already projected on the ray (atmosphere::project/transform,
 * atmos/atmos_rad.cc:15-51).  Layout state[col][field][z], depth index z=0 at
 * the top of the atmosphere (reference x3l) and z=nd-1 at the bottom (x3h),
 * nfield = SNAPI_NF_BASE + 2*nline.  Double for SNAPI_FP64, float for
 * SNAPI_MIXED.
 */
enum snapi_field {
  SNAPI_F_T = 0,     /* temperature [K] */
  SNAPI_F_VT = 1,    /* microturbulent velocity [cm/s] */
  SNAPI_F_VLOS = 2,  /* velocity projected on the ray [cm/s] */
  SNAPI_F_BMAG = 3,  /* |B| [G] */
  SNAPI_F_BINC = 4,  /* inclination of B to the ray [deg] */
  SNAPI_F_BAZI = 5,  /* azimuth of B about the ray [deg] */
  SNAPI_F_NE = 6,    /* electron density [cm^-3] */
  SNAPI_F_NHI = 7,   /* neutral hydrogen density [cm^-3] (H- and van der Waals damping) */
  SNAPI_F_OPREF = 8, /* referent continuum opacity at 500 nm [cm^-1] (TAU_GRID=1 normalisation) */
  SNAPI_F_GRID = 9,  /* RT grid: tau_referent (negative, -10^logtau) or height [cm] */
  SNAPI_F_BGOP = 10, /* wavelength-independent background opacity (Thomson + atomic bound-free) [cm^-1] */
  SNAPI_F_BGEM = 11, /* its emissivity */
  SNAPI_F_NT = 12,   /* total particle density [cm^-3] (density responses) */
  SNAPI_NF_BASE = 13 /* then per line l: SNAPI_NF_BASE+2l lower, +2l+1 upper level population */
};

#define SNAPI_MAX_LINES 16
#define SNAPI_MAX_COMPONENTS 96
#define SNAPI_MAX_ND 256
#define SNAPI_MAX_NODES 16

/* One bound-bound transition (atom::boundbound_op_em_vector,
 * atmos/atmol/atom/atom_rad_synthesis.cc:396-505). */
typedef struct {
  double lam0;   /* vacuum line centre h c / (E_u - E_l) [cm] */
  double aul;    /* A_ul [s^-1] */
  double blu;    /* B_lu in the reference's units (atom.cc:238-243) */
  double bul;    /* B_ul */
  double grad;   /* radiative damping: sum of A out of both levels (atom.cc:926-935) */
  double sigma;  /* van der Waals cross-section (atom::compute_damp_col) */
  double alpha;  /* its velocity exponent */
  double mass;   /* atomic mass [g] */
  int split;     /* 1: Zeeman split (atom.cc:1650-1651); 0: scalar Voigt */
  int ncomp[3];  /* number of pi, sigma_b, sigma_r components */
  int comp0;     /* offset of this line's components in the component table */
} snapi_line;

/* Zeeman components: shift in units of lambda_B = 4.67e-5 lam0^2 |B| and
 * relative strength (atom::zeeman_setup, atom.cc:1620-1773). */
typedef struct {
  double shift;
  double strength;
} snapi_component;

typedef struct {
  int nline;
  int ncomp_total;
  snapi_line line[SNAPI_MAX_LINES];
  snapi_component comp[SNAPI_MAX_COMPONENTS];
} snapi_line_table;

/*
 * Building the line table from level data, as the reference's atom class
 * does (host functions; no device is touched).
 *
 * snapi_line_add appends one line to t (and its Zeeman components to
 * t->comp): lam0 = h c / (E_u - E_l); B_ul = lam0^3 / (2 h c) A_ul and
 * B_lu = g_u / g_l B_ul, both times lam0^2 / c (atom::atom,
 * atmos/atmol/atom/atom.cc:233-253); the van der Waals cross-section and
 * exponent of snapi_vdw_damping; the Zeeman pattern of
 * snapi_zeeman_pattern.  stage is the ionisation stage of both levels (0
 * neutral), e_ion its ionisation energy [erg], energies [erg] from the
 * stage's ground level, grad the radiative damping (snapi_radiative_damping).
 * For hydrogen the reference sets l = level index - 1 (atom.cc:999-1004);
 * pass l that way.
 *
 * snapi_vdw_damping: atom::compute_damp_col (atom.cc:987-1308): effective
 * quantum numbers n* = (stage + 1) sqrt(13.6 eV / (e_ion - E)); for neutral
 * s-p, p-d and d-f transitions (l_lower + l_upper = 1, 3, 5) the
 * Anstee-Barklem-O'Mara tables interpolated at the two n* (interpol_2d,
 * cmlib/mathtools.cc:1006-1041), otherwise Unsold's C6 with alpha = 0.4.
 *
 * snapi_zeeman_pattern: atom::zeeman_setup (atom.cc:1620-1773) for levels of
 * 2J = j2_lower, j2_upper and Lande factors g_lower, g_upper: split = 1 for
 * Delta J = +-1 and for Delta J = 0 with J != 0; then ncomp[0] pi, ncomp[1]
 * sigma_b and ncomp[2] sigma_r components written to comp in that order,
 * shifts in units of lambda_B and strengths normalised per group.
 *
 * snapi_radiative_damping: atom::damp_rad (atom.cc:926-935), the sum of the
 * A of every transition out of both levels; A [nlevel][nlevel], A[u][l].
 */
typedef struct {
  double energy;  /* excitation energy [erg] */
  double g;       /* statistical weight */
  int j2;         /* 2 J */
  int l;          /* orbital quantum number of the level's configuration */
  double g_lande; /* Lande factor */
} snapi_level;
int snapi_zeeman_pattern(int j2_lower, int j2_upper, double g_lower, double g_upper, int *split, int *ncomp,
                         snapi_component *comp, int max_comp);
int snapi_vdw_damping(int stage, double e_ion, double e_lower, double e_upper, int l_lower, int l_upper, double *sigma,
                      double *alpha);
int snapi_radiative_damping(int nlevel, const double *A, int lower, int upper, double *grad);
int snapi_line_add(snapi_line_table *t, double mass, int stage, double e_ion, const snapi_level *lower,
                   const snapi_level *upper, double aul, double grad);

/* Ray and grid options shared by the synthesis and response entry points. */
typedef struct {
  double theta;    /* reference 'el' (pi - theta): pi for mu = 1 */
  double lam_mid;  /* wavelength of the Thomson emissivity, (lambda[1]+lambda[n])/2 (vacuum) [cm] */
  int tau_grid;    /* 1: op/em normalised by op_ref and GRID is tau_referent */
  int hminus;      /* 1: add H- bound-free + free-free (atmos/atmol/mol/mol.cc:379-412) */
  int hminus_mid;  /* 1: H- at lam_mid for every wavelength, as the response path of the
                      reference evaluates it (atmos/atmol/mol/mol.cc:414-469) */
} snapi_ray;

/* Library version / device query. */
int snapi_version(void);
int snapi_device_count(void);

/*
 * Launch-shape autotuner (no counterpart in the reference, whose solver runs
 * one column per CPU process).  The synthesis kernels come in three
 * occupancy bounds (float: 8, 7, 6 wavefronts per SIMD; double: the
 * compiler's own, 3, 4), the node-Jacobian tangent kernels in workgroups
 * of 64, 128 and 256 work-items, and the float ALI accumulation of atoms of
 * up to 10 lines and 6 bound-free transitions in bounds of 4, 3 and 2
 * wavefronts per SIMD.  Every shape computes the same results bit
 * for bit.  With tuning on (the default; SNAPI_AUTOTUNE=0 in the environment
 * turns it off), the first call for a problem shape (kernel, precision,
 * depths, lines, components, wavelengths, node pass) on a device times each
 * shape on the call's own first columns (up to 16384 of them) with HIP
 * events and keeps the fastest for the rest of the process; calls made
 * while their stream is being captured into a graph, and calls with tuning
 * off, take the first shape of the list unless the shape has been chosen
 * already.
 *   snapi_autotune_mode: 0 off, 1 on, -1 query; returns the previous mode.
 *   snapi_autotune_force: kernel 0 synthesis, 1 tangent, 2 ALI
 *     accumulation; candidate index to launch from now on, -1 to return to
 *     the tuned choice.
 *   snapi_autotune_clear: forgets every choice.
 *   snapi_autotune_report: one JSON object per line and choice (device,
 *     kernel, precision, nd, nline, ncomp, nlambda, extra, choice, ms of each
 *     candidate, -1 for one that failed) into buf of n bytes
 *     (SNAPI_ERR_CAPACITY when it does not fit).
 */
int snapi_autotune_mode(int mode);
int snapi_autotune_force(int kernel, int candidate);
int snapi_autotune_clear(void);
int snapi_autotune_report(char *buf, size_t n);

/*
 * Kernel status.  The kernels report what they cannot compute (a regime the
 * build does not implement, bit 1 -> SNAPI_ERR_UNSUPPORTED) and non-finite
 * results (bit 2 -> SNAPI_ERR_NONFINITE) by OR-ing bits into a status word on
 * the device.  The host-pointer entry points clear the word before their
 * launches and return what it holds after them.  The *_dev entry points only
 * launch, so their callers read the word when their stream's work is done.
 *
 * A caller owns a word by creating a context on the current device and
 * binding it to the calling thread: every launch that thread makes on that
 * device then reports into the context's word, and no other caller's calls
 * touch it.  Without a bound context the launches report into a word of the
 * device, which snapi_get_status reads.  No counterpart in the reference,
 * whose functions log and return an int (or exit) at the point of failure.
 *   snapi_context_create: a context (and its zeroed word) on the current device;
 *   snapi_context_bind: ctx for the calling thread's later calls (NULL: the
 *     device's word again);
 *   snapi_context_status / snapi_get_status: synchronise `stream` (a
 *     hipStream_t, null = the default stream), return SNAPI_OK,
 *     SNAPI_ERR_UNSUPPORTED or SNAPI_ERR_NONFINITE for the bits the word
 *     holds, and clear them when reset is nonzero.
 */
typedef struct snapi_context snapi_context;
int snapi_context_create(snapi_context **ctx);
int snapi_context_destroy(snapi_context *ctx);
int snapi_context_bind(snapi_context *ctx);
int snapi_context_status(snapi_context *ctx, void *stream, int reset);
int snapi_get_status(void *stream, int reset);

/*
 * Polarized synthesis of a batch of columns on one wavelength grid.
 *   state   [ncol][nfield][nd] (double for SNAPI_FP64, float for SNAPI_MIXED)
 *   lam     [nlambda] vacuum wavelengths [cm] (double)
 *   stokes  [ncol][nlambda][4] emergent I,Q,U,V (double for SNAPI_FP64, float for SNAPI_MIXED)
 * One work-item per (column, wavelength); the column tile is staged in LDS.
 */
int snapi_synth_stokes(int precision, const snapi_line_table *lines, const snapi_ray *ray, int ncol, int nd,
                       const void *state, int nlambda, const double *lam, void *stokes);
int snapi_synth_stokes_dev(int precision, const snapi_line_table *lines, const snapi_ray *ray, int ncol, int nd,
                           const void *state_dev, int nlambda, const double *lam_dev, void *stokes_dev,
                           void *stream);

/*
 * Per-depth responses d(IQUV)/d q_p(z) of the emergent Stokes vector along
 * the seven physical parameters of the reference: p = 0 T, 1 Nt, 2 vt,
 * 3 vLOS, 4 |B|, 5 inclination, 6 azimuth (angles per radian), computed by a
 * reverse (adjoint) sweep of the formal solution with the four emergent
 * Stokes components as seeds.  dstate holds the thermodynamic derivatives
 * the reference obtains by differencing its equation of state
 * (ne_lte_derivatives, LTE level responses, compute_op_referent_derivative):
 * [ncol][2][nfield][nd], slice 0 = d/dT, slice 1 = d/dNt; the fields NE,
 * NHI, OPREF and the populations are read.  For the reference's response
 * path set ray->hminus_mid = 1.
 *   stokes [ncol][nlambda][4], resp [ncol][7][nd][nlambda][4]
 * All arrays in the precision's type (double for SNAPI_FP64, float for
 * SNAPI_MIXED), lam in double.  Valid for the mu = 1 ray (theta = pi), where
 * the ray and atmosphere frames coincide (atmosphere::transform_responses,
 * atmos/atmos_rad.cc:53-101, is the identity there); any other ray returns
 * SNAPI_ERR_UNSUPPORTED, as does snapi_stokes_adjoint.  For other rays use
 * snapi_stokes_responses_atm.
 */
int snapi_stokes_responses(int precision, const snapi_line_table *lines, const snapi_ray *ray, int ncol, int nd,
                           const void *state, const void *dstate, int nlambda, const double *lam, void *stokes,
                           void *resp);

/*
 * Per-depth responses for any ray, in the atmosphere's frame, as
 * atmosphere::obs_stokes_responses returns them (atmos/atmos_rad_synthesis.cc:
 * 166-399, transform_responses at :344): the responses of
 * snapi_stokes_responses taken on the ray (theta = ray->theta, phi; the
 * state holds the ray-frame projections, as snapi_cube_to_state_dev writes
 * them), then transformed by snapi_transform_responses with the
 * atmosphere-frame field bvec [ncol][3][nd] (Bx, By, Bz in G, double).
 */
int snapi_stokes_responses_atm(int precision, const snapi_line_table *lines, const snapi_ray *ray, double phi,
                               const double *bvec, int ncol, int nd, const void *state, const void *dstate,
                               int nlambda, const double *lam, void *stokes, void *resp);
/*
 * atmosphere::transform_responses (atmos/atmos_rad.cc:53-101) on per-depth
 * responses resp [ncol][7][nd][nlambda][4] (precision's type), in place: per
 * depth, with the atmosphere-frame field bvec [ncol][3][nd] (double), the
 * inclination and azimuth responses become
 *   R_incl' = R_incl d(incl_ray)/d(incl_atm) + R_azi d(azi_ray)/d(incl_atm),
 *   R_azi'  = R_incl d(incl_ray)/d(azi_atm)  + R_azi d(azi_ray)/d(azi_atm),
 * with incl = acos(B_z/|B|) and azi = atan(B_y/B_x) in each frame (the ray
 * frame rotated by theta and phi as atmosphere::transform, :23-51), and the
 * velocity response is multiplied by cos(phi), as the reference writes it
 * (:89).  Where |B| = 0, the field lies along the ray or B_x = B_y = 0, where
 * the reference divides by zero, both angle responses are set to zero.
 */
int snapi_transform_responses(int precision, int ncol, int nd, int nlambda, const double *bvec, double theta,
                              double phi, void *resp);
int snapi_transform_responses_dev(int precision, int ncol, int nd, int nlambda, const double *bvec_dev, double theta,
                                  double phi, void *resp_dev, void *stream);

/*
 * Gradient of chi^2/2 = sum_{l,s} w_s (S - S_obs)^2 / 2 with respect to the
 * per-depth parameters in one reverse sweep (one seed w_s (S - S_obs)), i.e.
 * J^T r without forming J.
 *   obs [ncol][nlambda][4], grad [ncol][7][nd] (precision's type); wstokes[4] double
 */
int snapi_stokes_adjoint(int precision, const snapi_line_table *lines, const snapi_ray *ray, int ncol, int nd,
                         const void *state, const void *dstate, int nlambda, const double *lam, const void *obs,
                         const double *wstokes, void *grad);

/*
 * One damped normal-equation solve per pixel (batched):
 *   J [npix][nrow][npar] (row = (l-1)*4 + s as atmos_fit.cc:129-133), r [npix][nrow]
 *   JTJ[npix][npar][npar], rhs[npix][npar] are written (undamped J^T J, J^T r),
 *   corr[npix][npar] = pinv(JTJ * diag(1+lm)) rhs with singular values below
 *   w_max/1e6 dropped (solve_svd semantics), in double.
 * J^T J is accumulated on MFMA from float J (SNAPI_MIXED) or in double (SNAPI_FP64).
 */
int snapi_lm_step(int precision, int npix, int nrow, int npar, const void *J, const void *r, const double *lm,
                  double *JTJ, double *rhs, double *corr);

/* Node model of the inversion (the reference's model class, atmos/atmos.h:
 * 377-531): one entry per node, parameter code as the reference orders them
 * (1 T, 2 vt, 3 vLOS, 4 |B|, 5 inclination [rad], 6 azimuth [rad]) and the
 * node's position in log tau_500.
 * reg_type[p], reg_alpha[p] (p = parameter code, index 0 unused): the
 * model's regularisation of that parameter (model::temp_reg_type ... and
 * their alphas, set from REG_TYPE / REG_ALPHA of the parameter's block,
 * jsub/cfg.cc:374-381): 0 none, 1 departure from zero, 2 Tikhonov
 * (differences of neighbouring nodes).  The fits add alpha^2 R^T R of each
 * regularised parameter to J^T J before the damping
 * (atmosphere::regularize_hessian, atmos/atmos_fit.cc:1371-1480), the
 * right-hand side unchanged as there.  R covers the parameter's first run
 * of consecutive nodes; as in the reference, a run that ends the node list
 * contributes nothing to departure-from-zero and only its first difference
 * to Tikhonov, and a single Tikhonov node couples to the node after it.  A
 * parameter without nodes is not regularised; an entry of R beyond the
 * last node is dropped. */
typedef struct {
  int nnode;
  int param[SNAPI_MAX_NODES];
  double logtau[SNAPI_MAX_NODES];
  int reg_type[7];
  double reg_alpha[7];
} snapi_node_model;

/* Instrument effects the fits apply (atmosphere::stokes_lm_fit,
 * atmos/atmos_fit.cc:104-115, and forward_evaluate, :717-737); a null
 * pointer means none.  broadening [cm] > 0: the Gaussian convolution of
 * snapi_spectral_convolve; npsf > 0: the sampled PSF of snapi_psf_convolve
 * (psf in host memory, npsf <= 1024).  As the reference: the current
 * spectrum and the Jacobian take the Gaussian when broadening > 0 and
 * otherwise the PSF; every trial spectrum takes the Gaussian, then the
 * PSF. */
typedef struct {
  double broadening;
  int npsf;
  const double *psf;
} snapi_instrument;

enum snapi_jacobian {
  SNAPI_JAC_ADJOINT = 0,     /* reverse sweep contracted with the node responses */
  SNAPI_JAC_DIFFERENCES = 1, /* central differences of the node values through the synthesis (2n+1 syntheses) */
  SNAPI_JAC_TANGENT = 2,     /* forward-mode sweep carrying every node direction (no per-depth storage) */
  /* flag, ORed with one of the above: the node-to-depth maps as the exact
   * derivatives of the node interpolation and field projection, and, for
   * temperature nodes, dNt/dnode by the hydrostatic tangent
   * (snapi_hequilibrium_tangent), instead of the reference's central
   * differences of the interpolation and its two hydrostatic rebuilds per
   * temperature node (atmos/atmos_fit.cc:1225-1305) */
  SNAPI_JAC_ANALYTIC_MAPS = 16
};

/*
 * Batched Levenberg-Marquardt inversion (atmosphere::stokes_lm_fit,
 * atmos/atmos_fit.cc:18-227) of npix pixels for the kinematic and magnetic
 * node parameters (codes 2..6), the thermodynamic state of each pixel held
 * as given; temperature nodes are rejected (SNAPI_ERR_UNSUPPORTED).  mu = 1 ray.
 *   state [npix][nfield][nd], obs [npix][nlambda][4], mask [npix][nlambda] (or
 *   null: all 1) in the precision's type; nodes [npix][nnode] in: start
 *   values, out: solution; chi2 [npix], iterations [npix] out.
 * nslot pixels are iterated at once; a finished pixel's slot is refilled from
 * the queue of the remaining ones; nslot <= 0: one slot per pixel, capped at
 * what half of the free device memory holds.  The host reads the active
 * count only every k iterations (k doubling while a batch of iterations
 * takes under 2 ms, up to 32); the _dev variants on a stream other than the
 * null stream replay the iterations after the first from one captured HIP
 * graph (SNAPI_LAUNCH=eager in the environment: eager launches; the results
 * are the same bit for bit).  pixel_iterations (optional) returns the total
 * number of LM iterations over all pixels, counted on the device.
 */
int snapi_invert(int precision, int jacobian, const snapi_line_table *lines, const snapi_ray *ray,
                 const snapi_node_model *model, int npix, int nd, const void *state, int nlambda, const double *lam,
                 const void *obs, const void *mask, const double *wstokes, double lm_start, int max_iter, int nslot,
                 double *nodes, double *chi2, int *iterations, long long *pixel_iterations);

/*
 * Emergent Stokes vector and node Jacobian of npix pixels at given node
 * values (atmosphere::obs_stokes_responses_to_nodes, atmos/atmos_fit.cc:
 * 1205-1352, for node parameter codes 2..6; mu = 1 ray).  The node-controlled
 * fields of state are replaced by the node interpolation first.
 *   state [npix][nfield][nd] and the outputs in the precision's type;
 *   nodes [npix][nnode] double;
 *   stokes [npix][nlambda][4], J [npix][nlambda*4][nnode] = dS_s(lambda_l)/d node_n
 *   (per cm/s, G or rad; row l*4+s).
 * jacobian selects the method (snapi_jacobian).
 */
int snapi_node_jacobian(int precision, int jacobian, const snapi_line_table *lines, const snapi_ray *ray,
                        const snapi_node_model *model, int npix, int nd, const void *state, const double *nodes,
                        int nlambda, const double *lam, void *stokes, void *J);

/*
 * snapi_invert with the arrays already in device memory (state_dev, lam_dev,
 * obs_dev, mask_dev as above; nodes0_dev [npix][nnode] start values,
 * nodes_dev, chi2_dev, iterations_dev out), all work on `stream` (a
 * hipStream_t; null = the default stream).  Same semantics and errors.
 */
int snapi_invert_dev(int precision, int jacobian, const snapi_line_table *lines, const snapi_ray *ray,
                     const snapi_node_model *model, int npix, int nd, const void *state_dev, int nlambda,
                     const double *lam_dev, const void *obs_dev, const void *mask_dev, const double *wstokes,
                     double lm_start, int max_iter, int nslot, const double *nodes0_dev, double *nodes_dev,
                     double *chi2_dev, int *iterations_dev, long long *pixel_iterations, void *stream);

/*
 * Scalar formal solution with the diagonal approximate Lambda operator for a
 * batch of (column, frequency, ray) problems:
 *   op, em [nprob][nd], grid [nprob][nd]; I, lstar [nprob][nd]
 *   ct[nprob] = cos(theta) of each ray; boundary as the reference's (-1: S at the bottom).
 */
int snapi_formal_scalar(int nprob, int nd, const double *grid, const double *op, const double *em, const double *ct,
                        double boundary, double *I, double *lstar);

/* ---------------------------------------------------------------------------
 * ALI (approximate Lambda iteration) for one NLTE atom in plane-parallel
 * columns: atmosphere::nltepops (atmos/atmos_pop.cc:67-240) with
 * atom::add (atmos/atmol/atom/atom.cc:1865-1927) and atom::newpops
 * (atmos/atmol/atom/atom_pop.cc:27-135).
 *
 * Levels are the atom's mapped levels (atom::nmap order: every level of every
 * ionisation stage, the last one being the ground level of the final stage).
 * Transitions are the atom's own list (atom::tmap order): lines between levels
 * of one stage, and one bound-free transition from every level except the
 * last to the ground level of the next stage.
 */
#define SNAPI_ALI_MAX_LEVELS 16
#define SNAPI_ALI_MAX_TRANS 64
enum snapi_transition_kind { SNAPI_TR_LINE = 0, SNAPI_TR_BF = 1 };
typedef struct {
  int kind;      /* SNAPI_TR_LINE or SNAPI_TR_BF */
  int lower;     /* mapped level index (bound-free: the ionising level) */
  int upper;     /* mapped level index (bound-free: ground level of the next stage) */
  double lam0;   /* line: h c / (E_u - E_l) [cm] */
  double aul;    /* line: A_ul [s^-1] */
  double blu;    /* line: B_lu as the reference stores it (atom.cc:237-243) */
  double bul;    /* line: B_ul */
  double energy; /* line: E_u - E_l [erg] */
  double gratio; /* bound-free: g_lower / g_continuum */
  double de;     /* bound-free: ionisation energy of the stage minus E_lower [erg] */
} snapi_ali_transition;
typedef struct {
  int nlevel;
  int ntrans;
  double mass;   /* atomic mass [g] */
  snapi_ali_transition tr[SNAPI_ALI_MAX_TRANS];
} snapi_ali_atom;

/* Per-depth column fields of the ALI entry points: colstate[col][field][z]. */
enum snapi_ali_field {
  SNAPI_A_T = 0,     /* temperature [K] */
  SNAPI_A_VT = 1,    /* microturbulent velocity [cm/s] */
  SNAPI_A_NE = 2,    /* electron density [cm^-3] */
  SNAPI_A_OPREF = 3, /* referent opacity (normalisation of the formal solution when tau_grid) */
  SNAPI_A_GRID = 4,  /* RT grid (tau_referent, negative, or height) */
  SNAPI_A_NA = 5,    /* total population of the atom (conservation equation) */
  SNAPI_A_NF = 6
};

/* Frequency grid, rays and the fixed (population-independent) inputs. */
typedef struct {
  int nlam;               /* wavelength points */
  const double *lam;      /* [nlam] vacuum wavelengths [cm] */
  const double *lamw;     /* [nlam] integration weights */
  const double *fudge;    /* [nlam] opacity fudge of the bound-free terms */
  const double *sigma;    /* [ntrans][nlam] bound-free cross-sections (line rows unused) */
  int nray;
  const double *theta;    /* [nray] ray angles as the reference's anglesetup gives them */
  const double *wray;     /* [nray] angular weights (bin) */
  int bg_per_ray;         /* 1: bgop/bgem hold one set per ray; 0: one set for all rays */
  double boundary;        /* boundary condition of the formal solution (-1: S at the bottom) */
  int tau_grid;           /* 1: opacities normalised by SNAPI_A_OPREF in the formal solution */
} snapi_ali_grid;

/*
 * One ALI accumulation pass with the given populations: for every ray and
 * wavelength the opacity (background + the atom's bound-free and lines with
 * Voigt profiles), the scalar formal solution with Lambda*, and the sums of
 * atom::add.
 *   colstate [ncol][SNAPI_A_NF][nd], vproj [ncol][nray][nd] velocity along
 *   each ray, gamma [ncol][ntrans][nd] total damping of each line,
 *   bgop, bgem [ncol][nray or 1][nd][nlam] background opacity/emissivity,
 *   pops [ncol][nd][nlevel];
 *   acc [ncol][nd][ntrans][4] out: J (lower->upper), J (upper->lower),
 *   Lambda*, profile norm (atom::Jb, Ju, Ls, norm).  All double.
 * precision selects the formal solution's arithmetic; sums are in double.
 */
int snapi_ali_accumulate(int precision, const snapi_ali_atom *atom, const snapi_ali_grid *grid, int ncol, int nd,
                         const double *colstate, const double *vproj, const double *gamma, const double *bgop,
                         const double *bgem, const double *pops, double *acc);

/*
 * New populations from the accumulated sums (atom::newpops): rate matrix with
 * the local approximate operator (alo = 1) or plain Lambda iteration
 * (alo = 0), collision rates coll [ncol][nd][nlevel][nlevel] (rate i -> j),
 * the conservation equation replacing the row of the most populated level,
 * pivot-free Crout solve in double.
 *   newpops [ncol][nd][nlevel] out; relchange [ncol] out: max over depth and
 *   level of |new/old - 1|, or -1 where a population came out negative.
 */
int snapi_ali_newpops(const snapi_ali_atom *atom, int ncol, int nd, const double *colstate, const double *coll,
                      const double *pops, const double *acc, int alo, double *newpops, double *relchange);

/*
 * The whole iteration on the device (atmosphere::nltepops): accumulate,
 * new populations with the approximate operator, plain Lambda iteration for
 * a column whose populations came out negative, stop a column when its
 * relative change is below tol or after max_iter iterations.
 *   pops [ncol][nd][nlevel] in: start populations, out: result;
 *   iterations [ncol], relchange [ncol] (last value; -1: negative populations
 *   with both methods, the column keeps its previous populations) out.
 */
int snapi_ali_solve(int precision, const snapi_ali_atom *atom, const snapi_ali_grid *grid, int ncol, int nd,
                    const double *colstate, const double *vproj, const double *gamma, const double *bgop,
                    const double *bgem, const double *coll, double tol, int max_iter, double *pops, int *iterations,
                    double *relchange);

/* snapi_ali_solve on device-resident arrays (colstate, vproj, gamma, bgop,
 * bgem, coll, pops, iterations, relchange: device pointers; the grid's
 * arrays stay host arrays and are copied), all work on `stream`. */
int snapi_ali_solve_dev(int precision, const snapi_ali_atom *atom, const snapi_ali_grid *grid, int ncol, int nd,
                        const double *colstate_dev, const double *vproj_dev, const double *gamma_dev,
                        const double *bgop_dev, const double *bgem_dev, const double *coll_dev, double tol,
                        int max_iter, double *pops_dev, int *iterations_dev, double *relchange_dev, void *stream);

/*
 * snapi_ali_solve_dev without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  Same arguments,
 * semantics and kernels as snapi_ali_solve_dev; the grid's arrays, the sums
 * and the iteration's work areas live in a fixed device area of
 * snapi_ali_noalloc_capacity() bytes (64 MiB), one per device, held until the
 * process ends.  A problem needing more returns SNAPI_ERR_CAPACITY before any
 * device work (about 2,400 columns of 81 depths for an atom of 10
 * transitions and 220 wavelengths).  Every call on a device shares that
 * device's area and writes every value it reads from it, so calls that
 * overlap on one device (on different streams or threads) race; calls in
 * sequence do not.  It calls no hipMalloc, hipFree or allocating library
 * routine.
 */
size_t snapi_ali_noalloc_capacity(void);
int snapi_ali_solve_dev_noalloc(int precision, const snapi_ali_atom *atom, const snapi_ali_grid *grid, int ncol,
                                int nd, const double *colstate_dev, const double *vproj_dev,
                                const double *gamma_dev, const double *bgop_dev, const double *bgem_dev,
                                const double *coll_dev, double tol, int max_iter, double *pops_dev,
                                int *iterations_dev, double *relchange_dev, void *stream);

/*
 * Chemical equilibrium and LTE populations, per depth point, in double.
 *
 * The species table mirrors what the reference's atom and molecule objects
 * hold for atmosphere::chemeq: per element the number abundance, the
 * ionisation energies and one partition function per ionisation stage
 * (pf/cpf/trav/irwin, atmos/atmol/atom/partf.cc:14-358); per molecule its
 * components and its dissociation-constant form (eqc_cnst/eqc_poly/eqc_hmin,
 * atmos/atmol/mol/moleqc.cc); and the levels whose Boltzmann populations
 * (atom::lte, atmos/atmol/atom/atom.cc:779-787) are wanted.
 */
#define SNAPI_CHEM_MAX_ELEMENTS 14
#define SNAPI_CHEM_MAX_STAGES 10
#define SNAPI_CHEM_MAX_TERMS 6
#define SNAPI_CHEM_MAX_COEFS 8
#define SNAPI_CHEM_MAX_MOLECULES 32
#define SNAPI_CHEM_MAX_COMPONENTS 4
#define SNAPI_CHEM_MAX_COMP_STAGE 2 /* component stages 0 and 1 (after clamping at 0) */
#define SNAPI_CHEM_MAX_LEVELS 256
#define SNAPI_CHEM_MAX_BF 64

enum snapi_pf_type { SNAPI_PF_NONE = 0, SNAPI_PF_CONST = 1, SNAPI_PF_TRAV = 2, SNAPI_PF_IRWIN = 3 };
enum snapi_eqc_type { SNAPI_EQC_NONE = 0, SNAPI_EQC_CONST = 1, SNAPI_EQC_POLY = 2, SNAPI_EQC_HMIN = 3 };

typedef struct {
  int type;   /* snapi_pf_type; NONE: U = 1 */
  int nterm;  /* TRAV: asymptotic terms; IRWIN: polynomial coefficients */
  double value;  /* CONST: U */
  double charge; /* TRAV: effective charge seen by the outer electron (stage charge + 1) */
  double g0;     /* TRAV: statistical weight of the ground term */
  double g2[SNAPI_CHEM_MAX_TERMS];   /* TRAV per term: weight of its asymptotic sum */
  double xion[SNAPI_CHEM_MAX_TERMS]; /* series limit [eV] */
  double xl[SNAPI_CHEM_MAX_TERMS];   /* first principal quantum number of the asymptotic sum */
  int ncoef[SNAPI_CHEM_MAX_TERMS];   /* explicit levels of the term */
  double gamma[SNAPI_CHEM_MAX_TERMS][SNAPI_CHEM_MAX_COEFS]; /* their excitation energies [eV] */
  double alpha[SNAPI_CHEM_MAX_TERMS][SNAPI_CHEM_MAX_COEFS]; /* their statistical weights */
  double irwin[SNAPI_CHEM_MAX_COEFS]; /* IRWIN: ln U = sum_i irwin[i] (ln T)^i */
} snapi_partition;

typedef struct {
  int nstage;       /* ionisation stages (the reference's Z + 1) */
  double abundance; /* number abundance, any common scale */
  double eion[SNAPI_CHEM_MAX_STAGES]; /* ionisation energy of stage z -> z+1 [erg], z < nstage-1 */
  snapi_partition pf[SNAPI_CHEM_MAX_STAGES];
} snapi_chem_element;

typedef struct {
  int ncomp;
  int element[SNAPI_CHEM_MAX_COMPONENTS]; /* 0: electron; 1..nelement: element in table order */
  int stage[SNAPI_CHEM_MAX_COMPONENTS];   /* charge index as the reference stores it (0 neutral,
                                             1 singly ionised, -1 the electron); the molecule's
                                             charge is their sum */
  int eqc;   /* snapi_eqc_type */
  int ncoef; /* POLY: log10 K = sum_i coef[i] (5040/T)^i; CONST: K = coef[0] */
  double coef[SNAPI_CHEM_MAX_COEFS];
} snapi_chem_molecule;

typedef struct {
  int element, stage; /* element 1..nelement */
  double g;           /* statistical weight; 0 marks a placeholder whose population is the stage's */
  double energy;      /* excitation energy [erg] */
} snapi_chem_level;

/* A bound-free transition of the continuum opacity at the referent
 * wavelength (atom::opacity_continuum, atmos/atmol/atom/atom_rad.cc:151-187):
 * from a level of stage `stage` to the ground level of stage+1. */
typedef struct {
  int element, stage;     /* element 1..nelement */
  double g, energy;       /* lower level: statistical weight, excitation energy [erg] */
  double g_up, energy_up; /* ground level of the next stage (g_up = 0: a placeholder whose
                             population is the stage's) */
  double sigma;           /* cross-section at lambda_ref [cm^2] */
} snapi_chem_bf;

typedef struct {
  int nelement, nmolecule, nlevel, nbf;
  int hminus;        /* 1: the referent opacity includes H- bound-free + free-free of
                        element 1's neutral stage (h_minus_mol::opacity_continuum) */
  double lambda_ref; /* referent wavelength [cm] */
  snapi_chem_element element[SNAPI_CHEM_MAX_ELEMENTS];
  snapi_chem_molecule molecule[SNAPI_CHEM_MAX_MOLECULES];
  snapi_chem_level level[SNAPI_CHEM_MAX_LEVELS];
  snapi_chem_bf bf[SNAPI_CHEM_MAX_BF];
} snapi_chem_table;

/* Where the chemical-equilibrium results go in a column state (see
 * snapi_eos_state). */
typedef struct {
  int nline;                           /* lines of the line table: nfield = SNAPI_NF_BASE + 2 nline */
  int line_level[SNAPI_MAX_LINES][2];  /* table level (0-based) of each line's lower, upper level */
  int nhi_element;                     /* element (1-based) whose neutral-stage density is NHI */
  double lambda_mid;                   /* Thomson emissivity wavelength [cm] (vacuum) */
  double lambda_bg;                    /* wavelength of the bound-free background [cm] (vacuum) */
  double sigma_bg[SNAPI_CHEM_MAX_BF];  /* cross-section of table bf transition j at lambda_bg, 0: none */
} snapi_state_map;

/*
 * Chemical equilibrium (atmosphere::chemeq, atmos/atmos_chm.cc:28-363) at
 * every depth point of ncol columns, then LTE level populations (atom::lte,
 * atom.cc:779-787), as atmosphere::ltepops (atmos/atmos_pop.cc:30-45) runs
 * them.  Newton iteration on (Ne, N_1..N_nelement) with ionisation fractions
 * from atom::ionfrc (atom.cc:892-924) and molecules in dissociation
 * equilibrium; at most the reference's 51 Newton steps; converged when every
 * |defect/N| <= 1e-4.  The reference's Ne input only enters its dissociation
 * constants, none of whose forms depend on it, so there is none here.
 *   T, Nt [ncol][nd] in (K, cm^-3);
 *   ne [ncol][nd] out;
 *   stage [ncol][nstage_total][nd] out (may be NULL): N of every ionisation
 *     stage, elements in table order, stages ascending;
 *   molecule [ncol][nmolecule][nd] out (may be NULL);
 *   level [ncol][nlevel][nd] out (may be NULL);
 *   opref [ncol][nd] out (may be NULL): continuum opacity at lambda_ref
 *     (atmosphere::opacity_continuum, atmos/atmos_rad.cc:286-294): Thomson,
 *     the table's bound-free transitions with stimulated emission, H- if
 *     hminus; the referent opacity of compute_op_referent
 *     (plpar/atmos_ppbez_rts.cc:268-281) for lambda_ref = airtovac(5000 A);
 *   iterations [ncol][nd] out (may be NULL): Newton steps taken, -1 where the
 *     linear system was singular (the reference exits there).
 */
int snapi_chemeq(const snapi_chem_table *table, int ncol, int nd, const double *T, const double *Nt, double *ne,
                 double *stage, double *molecule, double *level, double *opref, int *iterations);

/*
 * NLTE population responses: the two numerical kernels of
 * atom::compute_nlte_population_responses (atmos/atmol/atom/atom_responses.cc:
 * 26-204) as atmosphere::compute_nlte_population_responses
 * (atmos/atmos_responses.cc:13-157) drives them.
 *
 * snapi_formal_responses: atmos_ppbez::formal_with_responses
 * (atmos/plpar/atmos_ppbez_rts.cc:349-778), the scalar formal solution of one
 * ray with the response of every intensity to every opacity and emissivity,
 * with the reference's linear weights (its Bezier branch needs a forward step
 * above 1e10) and Taylor weights below a step of 1e-2.
 *   grid [ncol][nd]; op, em [ncol][nprob][nd], normalised as the reference
 *   passes them; theta [nprob] (ray angle of each problem, radians);
 *   boundary: -1 (I = S at the bottom) or a fixed intensity;
 *   I, L [ncol][nprob][nd] out (intensity, Lambda*);
 *   Rop, Rem [ncol][nprob][nd][nd] out: dI_z/dop_z', dI_z/dem_z'.
 * 2 <= nd <= 256.
 */
int snapi_formal_responses(int ncol, int nprob, int nd, const double *grid, const double *op, const double *em,
                           const double *theta, double boundary, double *I, double *L, double *Rop, double *Rem);
int snapi_formal_responses_dev(int ncol, int nprob, int nd, const double *grid_dev, const double *op_dev,
                               const double *em_dev, const double *theta_dev, double boundary, double *I_dev,
                               double *L_dev, double *Rop_dev, double *Rem_dev, void *stream);

/*
 * snapi_response_solve: the Crout factorisation without pivoting and its
 * solves (cmlib/mathtools.cc:487-514) that turn the assembled response
 * matrix and right-hand sides into level responses
 * (atom_responses.cc:171-200), for nsys independent systems, in double with
 * the reference's operation order.
 *   M [nsys][n][n]; B, X [nsys][nrhs][n] (one right-hand side per row).
 */
int snapi_response_solve(int nsys, int n, int nrhs, const double *M, const double *B, double *X);

/*
 * The NLTE population-response system of one NLTE atom:
 * atmosphere::compute_nlte_population_responses (atmos/atmos_responses.cc:
 * 13-157) with, per ray, atom::compute_profile_norm (atmos/atmol/atom/
 * atom_rad.cc:885-930), per ray and wavelength the scalar formal solution
 * with responses and atom::add_response_contributions_new
 * (atom_responses.cc:668-1128), then the rate part of
 * atom::compute_nlte_population_responses (atom_responses.cc:26-170).
 *
 * The atom: levels in map order, grouped by ion stage with consecutive
 * indices within a stage; the top stage holds one level, the last.  Every
 * pair of levels of one stage with nonzero Einstein coefficients is a line
 * transition; each bound-free transition goes from a level of stage z to the
 * ground level of stage z+1.
 */
typedef struct {
  int nmap;                   /* levels */
  int ntr;                    /* transitions */
  int nstage;                 /* ion stages 0..nstage-1 (the reference's Z + 1) */
  int damping;                /* turn_on_damping: collisional damping in the profiles (0/1) */
  double mass;                /* atom mass [g] */
  const int *level_stage;     /* [nmap] ion stage of each level */
  const int *level_index;     /* [nmap] index of the level within its stage */
  const double *level_energy; /* [nmap] excitation energy [erg] */
  const double *level_g;      /* [nmap] statistical weight */
  const double *stage_ip;     /* [nstage] ionisation energy of each stage [erg] */
  const int *tr_type;         /* [ntr] SNAPI_TR_LINE or SNAPI_TR_BF */
  const int *tr_lower;        /* [ntr] map index of the lower level */
  const int *tr_upper;        /* [ntr] map index of the upper level (bound-free: next stage's ground) */
  const double *tr_A;         /* [ntr] A_ul of the lines [s^-1] */
  const double *tr_Blu;       /* [ntr] B_lu of the lines */
  const double *tr_Bul;       /* [ntr] B_ul of the lines */
  const double *tr_damp_rad;  /* [ntr] radiative damping of the lines (atom::damp_rad) */
  const double *tr_lambda_crit; /* [ntr] edge wavelength of the bound-free transitions [cm] */
} snapi_nlte_atom;

/* Fields of the per-depth input of the response system, [SNAPI_NLTE_NFIELD][nd]
 * per column: temperature [K], microturbulence [cm/s], electron density,
 * total particle density Nt, referent opacity, the formal solver's grid
 * (rt_grid), the electron-density derivatives d ne/dT and d ne/dNt
 * (atmosphere::get_ne_lte_derivative 1, 2), the referent-opacity derivatives
 * d op_ref/dT, /dNt, /dvt (get_op_referent_der 1-3, local), and the
 * derivatives of the atom's total population d N/dT, d N/dNt
 * (atom::derivative_active_population[_density]). */
enum {
  SNAPI_NLTE_T = 0, SNAPI_NLTE_VT, SNAPI_NLTE_NE, SNAPI_NLTE_NT, SNAPI_NLTE_OPREF, SNAPI_NLTE_GRID,
  SNAPI_NLTE_DNE_DT, SNAPI_NLTE_DNE_DNT, SNAPI_NLTE_DOPREF_DT, SNAPI_NLTE_DOPREF_DNT, SNAPI_NLTE_DOPREF_DVT,
  SNAPI_NLTE_DNACT_DT, SNAPI_NLTE_DNACT_DNT, SNAPI_NLTE_NFIELD
};

/*
 * snapi_nlte_response_system: the system (M, beta) of ncol columns, and
 * optionally its solution, the level responses.
 *   atom: as above; tau_grid: 1 when the formal solver runs on the tau grid
 *     (op and em normalised by the referent opacity, whose derivatives enter);
 *   boundary: lower boundary of the formal solution (-1: I = S);
 *   ray_theta, ray_weight [nray]: the angular quadrature (anglesetup);
 *   lambda, lambda_weight [nlam]: the NLTE wavelength grid and its
 *     trapezoidal weights;
 *   sigma [ntr][nlam]: bound-free cross-section of each bound-free transition
 *     at each wavelength (the bf object's U(lambda); rows of lines unused);
 *   depth [ncol][SNAPI_NLTE_NFIELD][nd];
 *   pops [ncol][nmap][nd]: level populations;
 *   damp [ncol][2][ntr][nd]: collisional damping of each line
 *     (atom::damp_col) and its temperature derivative (damp_col_der_T);
 *   coll [ncol][3][nmap][nmap][nd]: collisional rate from level i to j
 *     (atom::collisional_rates) and its full derivatives to T and Nt
 *     (derivative_collisions_full_temp / _density);
 *   vr [ncol][nray][nd]: line-of-sight velocity of each ray (project);
 *   opem [ncol][nray][nlam][8][nd]: total opacity, emissivity, and the LTE
 *     perturbations of opacity (T, Nt, vt) and emissivity (T, Nt, vt) at each
 *     depth (opacity_pert_lte / emissivity_pert_lte, local);
 *   M [ncol][n][n], beta [ncol][3][nd][n] out, n = nd * nmap: response_matrix
 *     and beta_Temp / beta_density / beta_v_micro (one right-hand side per
 *     perturbed depth);
 *   resp [ncol][3][n][nd] out (may be NULL): level_responses, the solution
 *     for every right-hand side (snapi_response_solve of M and beta).
 * 2 <= nd <= 256, nmap <= 8, ntr <= 32, nray <= 16, nray * nlam <= 65535;
 * SNAPI_ERR_UNSUPPORTED for atoms outside the layout above.  Double
 * throughout.  The work area (about 128 MB per column of a 6-level,
 * 15-transition atom on 2 x 444 problems) is allocated per call and
 * covers as many columns per pass as half the free device memory holds.
 */
int snapi_nlte_response_system(const snapi_nlte_atom *atom, int ncol, int nd, int nray, int nlam, int tau_grid,
                               double boundary, const double *ray_theta, const double *ray_weight,
                               const double *lambda, const double *lambda_weight, const double *sigma,
                               const double *depth, const double *pops, const double *damp, const double *coll,
                               const double *vr, const double *opem, double *M, double *beta, double *resp);
/* The same on device arrays (every array argument except the atom's tables is
 * a device pointer), enqueued on stream. */
int snapi_nlte_response_system_dev(const snapi_nlte_atom *atom, int ncol, int nd, int nray, int nlam, int tau_grid,
                                   double boundary, const double *ray_theta, const double *ray_weight,
                                   const double *lambda_dev, const double *lambda_weight_dev,
                                   const double *sigma_dev, const double *depth_dev, const double *pops_dev,
                                   const double *damp_dev, const double *coll_dev, const double *vr_dev,
                                   const double *opem_dev, double *M_dev, double *beta_dev, double *resp_dev,
                                   void *stream);

/*
 * Gaussian spectral convolution (convolve_spectra_with_gauss,
 * cmlib/mathtools.cc:1350-1427), as the fit applies it to the spectrum and,
 * component by component, to the node Jacobian (convolve_response_with_gauss,
 * :1473-1498; atmos_fit.cc:104-110): each series is resampled with
 * interpol_1d onto an even grid of step width/2 padded by 10 widths,
 * convolved with exp(-d^2/width^2)/(sqrt(pi) width) by the trapezoidal rule
 * and resampled back, in double.
 *   data [nseries][nlambda][ncomp] in place (double for SNAPI_FP64, float for
 *     SNAPI_MIXED): Stokes [npix][nlambda][4] (ncomp 4) or J
 *     [npix][nlambda*4][nnode] (ncomp 4 nnode);
 *   lam [nlambda] ascending, host memory (the grid is planned on the host);
 *   width [cm] > 0.  SNAPI_ERR_UNSUPPORTED when even one component's fine grid
 *   does not fit the workgroup's 60 KiB of LDS (about 2,500 fine points).
 */
int snapi_spectral_convolve(int precision, int nseries, int nlambda, int ncomp, const double *lam, double width,
                            void *data);
/* The same on a device array; lam in host memory, lam_dev its device copy. */
int snapi_spectral_convolve_dev(int precision, int nseries, int nlambda, int ncomp, const double *lam,
                                const double *lam_dev, double width, void *data_dev, void *stream);

/*
 * Convolution with a sampled spectral PSF (convolve_spectra_with_psf,
 * cmlib/mathtools.cc:1429-1465), as the fit applies it to the spectrum and
 * to the node Jacobian (convolve_response_with_psf, :1500-1522;
 * atmos/atmos_fit.cc:111-115, 730-734): the npsf samples psf[0..npsf-1],
 * on the wavelength grid's step, centred on each point,
 *   out[l] = sum_j in[l - (npsf-1)/2 + j] psf[j] / sum(psf),
 * over the samples that fall on the grid (no renormalisation at the ends),
 * in double.  Deviation: the reference reads its 1-based PSF array from
 * index 0 while normalising over 1..n, and near the first wavelength it
 * anchors the PSF at the clipped start instead of centring it; here every
 * point takes the n given samples centred on it.  In the interior the two
 * agree when the reference's samples are the given ones shifted by one.
 *   data [nseries][nlambda][ncomp] in place (double for SNAPI_FP64, float
 *   for SNAPI_MIXED), as for snapi_spectral_convolve; npsf 1..1024.
 */
int snapi_psf_convolve(int precision, int nseries, int nlambda, int ncomp, int npsf, const double *psf, void *data);
/* The same on device arrays (psf_dev in device memory). */
int snapi_psf_convolve_dev(int precision, int nseries, int nlambda, int ncomp, int npsf, const double *psf_dev,
                           void *data_dev, void *stream);

/*
 * Scattered light of observable::add_scattered_light (cmlib/obs.cc:412-428):
 * Stokes I of every wavelength raised by fraction * continuum_level.
 *   stokes [nseries][nlambda][4] in place (double for SNAPI_FP64, float for
 *   SNAPI_MIXED).
 */
int snapi_add_scattered_light(int precision, int nseries, int nlambda, double fraction, double continuum_level,
                              void *stokes);
int snapi_add_scattered_light_dev(int precision, int nseries, int nlambda, double fraction, double continuum_level,
                                  void *stokes_dev, void *stream);

/*
 * Cube I/O.  ANA f0 files, uncompressed (ana_fzread / ana_fzwrite,
 * cmlib/ana_read.cc:74-160, ana_write.cc:120-171): header blocks of 512
 * bytes, dims[0] fastest; data of type 0 int8, 1 int16, 2 int32, 3 float32,
 * 4 float64, 5 int64, converted to this (little-endian) host's byte order on
 * reading; written little-endian with one header block per 255 characters of
 * header text.  Crunch-compressed files give SNAPI_ERR_UNSUPPORTED.
 *   snapi_ana_info: type, ndim, dims[16] and whether compressed;
 *   snapi_ana_read: the data into a buffer of exactly the data's bytes.
 */
int snapi_ana_info(const char *path, int *type, int *ndim, int *dims, int *compressed);
int snapi_ana_read(const char *path, void *data, size_t bytes);
int snapi_ana_write(const char *path, const void *data, int type, int ndim, const int *dims, const char *header);

/*
 * The 12 x NX x NY x NZ model cube (atmosphere::read_spinor3d,
 * atmos/atmos_fio.cc:79-127; NZ fastest, as the ANA file (nz, ny, nx, 12)
 * holds it) to column states (col = x * ny + y): per point log tau, height,
 * T, Pg, Pe, rho, kappa_ref, |B|, vmicro, vz, inclination, azimuth [rad] give
 * GRID = -10^log tau, T (at least 3200 K), NT = Pg / kT, NE = Pe / kT, OPREF,
 * VT, and the ray projections of the velocity and field for the ray (theta,
 * phi) (atmosphere::project / transform, atmos/atmos_rad.cc:15-51); the other
 * fields are zero, for snapi_eos_state to fill as popsetup and
 * compute_op_referent would.  cube_dev double; state in the precision's type
 * [nx*ny][nfield][nz].
 */
int snapi_cube_to_state_dev(int precision, const double *cube_dev, int nx, int ny, int nz, int nfield,
                            double theta, double phi, void *state_dev, void *stream);
/* Column Stokes spectra [nx*ny][nlambda][4] to the job's output cube
 * [nx][ny][4][nlambda] in double (master/job.cc:590-597; written as an ANA
 * FLOAT64 file with dims (nlambda, 4, ny, nx)). */
int snapi_stokes_to_cube_dev(int precision, const void *stokes_dev, int nx, int ny, int nlambda, double *cube_dev,
                             void *stream);

/*
 * Optical-depth scale on a height grid (atmos_ppbez::optical_depth_scale,
 * atmos/plpar/atmos_ppbez_rts.cc:221-266): per column the Bezier integral of
 * the opacity along the ray theta (reference convention: theta = 0 runs from
 * the top down) with Fritsch-Butland derivatives, one-sided at the ends, and
 * tau = 1e-8 at the first point.  With referent = 1 and theta = 0 this is
 * compute_tau_referent (:328-338): tau_referent = -tau from the referent
 * opacity, the grid the synthesis takes with tau_grid = 1.
 *   height [ncol][nd] (x3, cm; top first), op [ncol][nd] (cm^-1), tau [ncol][nd] out.
 */
int snapi_optical_depth_scale(int ncol, int nd, const double *height, const double *op, double theta, int referent,
                              double *tau);
int snapi_optical_depth_scale_dev(int ncol, int nd, const double *height_dev, const double *op_dev, double theta,
                                  int referent, double *tau_dev, void *stream);

/* snapi_chemeq on device-resident arrays (T, Nt and the outputs: device
 * pointers; the table stays a host struct and is copied), all work on
 * `stream`. */
int snapi_chemeq_dev(const snapi_chem_table *table, int ncol, int nd, const double *T_dev, const double *Nt_dev,
                     double *ne_dev, double *stage_dev, double *molecule_dev, double *level_dev, double *opref_dev,
                     int *iterations_dev, void *stream);

/*
 * Hydrostatic equilibrium: atmos_pp::enforce_hequilibrium
 * (atmos/plpar/atmos_pp.cc:372-455) for a batch of columns on their
 * referent optical-depth grids, top first.  The total particle density of
 * the top point is iterated from a 0.3 dyn/cm^2 guess on
 * Nt = 10^logtau rho g / kappa / kT; every deeper point takes the pressure
 * step g / kappa_mean * dtau (kappa_mean: geometric mean of the two points'
 * opacity per unit mass), iterated with chemical equilibrium at each new Nt
 * up to 20 times and until a correction is below 1e-2 of the top point's Nt;
 * g = 274.88e2 cm/s^2 (cmlib/const.h:25), rho = 1.4 * 1.67e-24 g times the
 * total density of element 1.  The outputs are chemical equilibrium at the
 * final Nt, as the reference's last pass (atmos_pp.cc:449-454) leaves them.
 * The opacity is taken at the table's lambda_ref, which must be set
 * (enforce_hequilibrium uses 5000 A as it stands, without airtovac).
 *   T, logtau [ncol][nd] in: temperature [K], log10 of the referent optical depth;
 *   Nt [ncol][nd] out: total particle density [cm^-3];
 *   ne, opref, rho [ncol][nd] out (may be NULL): electron density, opacity at
 *     lambda_ref [cm^-1], mass density [g cm^-3];
 *   stage, molecule, level out (may be NULL): as snapi_chemeq;
 *   iterations [ncol][nd] out (may be NULL): iterations at each depth (1..20).
 */
int snapi_hequilibrium(const snapi_chem_table *table, int ncol, int nd, const double *T, const double *logtau,
                       double *Nt, double *ne, double *opref, double *rho, double *stage, double *molecule,
                       double *level, int *iterations);
/*
 * The equation-of-state fields of a column state (double), as the reference's
 * synthesis path holds them after popsetup and compute_op_referent, and as
 * its response path holds them (obs_stokes_responses,
 * atmos/atmos_rad_synthesis.cc:173-190) with the thermodynamic derivatives
 * that snapi_stokes_responses and snapi_stokes_adjoint read:
 *   state [ncol][nfield][nd] (nfield = SNAPI_NF_BASE + 2 map->nline): T and
 *     NT in; NE, NHI (the neutral-stage density of map->nhi_element), OPREF
 *     (at the table's lambda_ref: airtovac(5000 A) for the reference's
 *     response and synthesis paths), BGOP and BGEM (Thomson with its
 *     emissivity at lambda_mid plus the map's bound-free transitions at
 *     lambda_bg, atom::boundfree_op_em_vector, atom_rad_synthesis.cc:226-262)
 *     and the line level populations out; other fields untouched;
 *   resp [ncol][nfield][nd] out (may be NULL): the response path's state:
 *     ne_lte_derivatives (atmos/atmos_responses.cc:364-423) leaves T raised
 *     by delta_T/2 = 2.5 K with Ne and the stage populations re-solved there
 *     and the level populations and referent opacity of T;
 *   dstate [ncol][2][nfield][nd] out (may be NULL): d/dT and d/dNt of NE, NHI,
 *     OPREF (compute_op_referent_derivative, plpar/atmos_ppbez_rts.cc:283-325),
 *     and of the line level populations
 *     (atom::compute_lte_population_responses_analytical,
 *     atmol/atom/atom_responses.cc:453-501); zero elsewhere.
 * Ten chemical equilibria per depth point with resp or dstate, one without.
 */
int snapi_eos_state(const snapi_chem_table *table, const snapi_state_map *map, int ncol, int nd, double *state,
                    double *resp, double *dstate);
int snapi_eos_state_dev(const snapi_chem_table *table, const snapi_state_map *map, int ncol, int nd,
                        double *state_dev, double *resp_dev, double *dstate_dev, void *stream);

/*
 * snapi_node_jacobian with temperature nodes (code 1), as
 * obs_stokes_responses_to_nodes (atmos_fit.cc:1205-1352) runs them: the
 * state's thermodynamics are rebuilt from the nodes (build_from_nodes,
 * atmos/plpar/atmos_pp.cc:58-180: the temperature interpolation on the
 * state's tau grid, hydrostatic equilibrium with chem_he, whose lambda_ref
 * is the 5000 A the march uses), the response path's state and
 * thermodynamic derivatives follow with chem and map (snapi_eos_state;
 * chem's lambda_ref airtovac(5000 A)), and the per-depth responses to T and
 * Nt are contracted with the temperature nodes' responses dT/dnode and
 * dNt/dnode, differences over delta_T = 5 K of two rebuilds each.  The
 * Stokes vector returned is the response path's, as the reference returns
 * it, with H- at ray->lam_mid as that path evaluates it
 * (atmos/atmol/mol/mol.cc:414-469) whatever ray->hminus_mid says.  At least
 * one temperature node.  SNAPI_JAC_ADJOINT contracts the per-depth T and Nt
 * responses with the node maps; SNAPI_JAC_TANGENT carries each temperature
 * node's direction dT/dnode d/dT + dNt/dnode d/dNt through the forward
 * sweep; SNAPI_JAC_DIFFERENCES is SNAPI_ERR_UNSUPPORTED here (the fits take
 * it, below).  J per K for temperature nodes.
 */
/*
 * snapi_invert[_dev] with temperature nodes: per pixel and iteration the
 * thermodynamics are rebuilt from the nodes as snapi_node_jacobian_eos
 * does; the Jacobian and the current chi^2 come from the response path, the
 * trial's chi^2 from the synthesis path's state (forward_evaluate), as
 * stokes_lm_fit takes them (atmos_fit.cc:18-227).  The state's
 * thermodynamic fields are rebuilt; its grid and, for parameters without
 * nodes, the fields stay as given.  As the reference: H- at ray->lam_mid on
 * the response path and at every wavelength on the trials (mol.cc:414-469,
 * 379-412) whatever ray->hminus_mid says; the start nodes bracketed into
 * the parameter limits before the first iteration (atmos_fit.cc:91); the
 * returned inclination and azimuth polished to acos(cos) and atan(tan)
 * (model::polish_angles, atmos/atmos_model.cc:992-1001).  At least one
 * temperature node.  Jacobians: SNAPI_JAC_ADJOINT and SNAPI_JAC_TANGENT as
 * snapi_node_jacobian_eos; SNAPI_JAC_DIFFERENCES differences the synthesis
 * path's spectrum over node sets each rebuilt through hydrostatic
 * equilibrium and the equation of state (the FP64 baseline; refused with
 * instrument effects).  inst (or null): the instrument effects of
 * snapi_instrument, applied to the current spectrum and Jacobian before the
 * Stokes weights and the mask, and to every trial spectrum.
 */
int snapi_invert_eos(int precision, int jacobian, const snapi_line_table *lines, const snapi_ray *ray,
                     const snapi_node_model *model, const snapi_chem_table *chem, const snapi_chem_table *chem_he,
                     const snapi_state_map *map, int npix, int nd, const void *state, int nlambda,
                     const double *lam, const void *obs, const void *mask, const double *wstokes,
                     const snapi_instrument *inst, double lm_start, int max_iter, int nslot, double *nodes,
                     double *chi2, int *iterations, long long *pixel_iterations);
int snapi_invert_eos_dev(int precision, int jacobian, const snapi_line_table *lines, const snapi_ray *ray,
                         const snapi_node_model *model, const snapi_chem_table *chem,
                         const snapi_chem_table *chem_he, const snapi_state_map *map, int npix, int nd,
                         const void *state_dev, int nlambda, const double *lam_dev, const void *obs_dev,
                         const void *mask_dev, const double *wstokes, const snapi_instrument *inst,
                         double lm_start, int max_iter, int nslot, const double *nodes0_dev, double *nodes_dev,
                         double *chi2_dev, int *iterations_dev, long long *pixel_iterations, void *stream);

/*
 * atmos_pp::build_from_nodes (atmos/plpar/atmos_pp.cc:58-180) followed by
 * popsetup and compute_op_referent, for npix pixels: the node-controlled
 * fields of state (double [npix][nfield][nd], nfield from map) from the
 * node values (nodes [npix][nnode]), T by the temperature interpolation on
 * the state's tau grid (GRID in), NT by hydrostatic equilibrium with
 * chem_he, and the equation-of-state fields of the synthesis path with
 * chem and map (snapi_eos_state).  At least one temperature node.
 */
int snapi_build_from_nodes(const snapi_node_model *model, const snapi_chem_table *chem,
                           const snapi_chem_table *chem_he, const snapi_state_map *map, int npix, int nd,
                           double *state, const double *nodes);
int snapi_build_from_nodes_dev(const snapi_node_model *model, const snapi_chem_table *chem,
                               const snapi_chem_table *chem_he, const snapi_state_map *map, int npix, int nd,
                               double *state_dev, const double *nodes_dev, void *stream);

int snapi_node_jacobian_eos(int precision, int jacobian, const snapi_line_table *lines, const snapi_ray *ray,
                            const snapi_node_model *model, const snapi_chem_table *chem,
                            const snapi_chem_table *chem_he, const snapi_state_map *map, int npix, int nd,
                            const void *state, const double *nodes, int nlambda, const double *lam, void *stokes,
                            void *J);

/* snapi_hequilibrium on device-resident arrays, all work on `stream`. */
int snapi_hequilibrium_dev(const snapi_chem_table *table, int ncol, int nd, const double *T_dev,
                           const double *logtau_dev, double *Nt_dev, double *ne_dev, double *opref_dev,
                           double *rho_dev, double *stage_dev, double *molecule_dev, double *level_dev,
                           int *iterations_dev, void *stream);

/*
 * Tangent of the hydrostatic equilibrium: for ncol columns already in
 * equilibrium (T, logtau, Nt [ncol][nd], Nt as snapi_hequilibrium returns
 * it), dNt [ncol][ndir][nd] along the temperature perturbations dT
 * [ncol][ndir][nd].  The linearised equations of the converged march
 * (P_0 = tau_0 g rho_0 / kappa_0; P_z - P_(z-1) = g dtau / kappa_mean), with
 * the density and opacity derivatives by central differences of the chemical
 * equilibrium (delta_T/2, delta_Nt_frac/2).  No counterpart in the
 * reference, whose obs_stokes_responses_to_nodes (atmos/atmos_fit.cc:
 * 1225-1248) differences two hydrostatic rebuilds per temperature node; this
 * replaces those rebuilds (SNAPI_JAC_ANALYTIC_MAPS).
 */
int snapi_hequilibrium_tangent(const snapi_chem_table *table, int ncol, int nd, const double *T,
                               const double *logtau, const double *Nt, int ndir, const double *dT, double *dNt);
int snapi_hequilibrium_tangent_dev(const snapi_chem_table *table, int ncol, int nd, const double *T_dev,
                                   const double *logtau_dev, const double *Nt_dev, int ndir, const double *dT_dev,
                                   double *dNt_dev, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* SNAPI_GPU_H */
