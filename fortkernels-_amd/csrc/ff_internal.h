// ff_internal.h — shared between the FF kernels (ff_device.hip) and the FF
// host side (ff_api.cpp).  Not public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace gpp {
namespace ff {

constexpr int kMaxJobBands = 8;   // bands sharing one pass over a slice
constexpr int kSliceThreads = 256;
constexpr int kSliceG = 2;        // G columns per slice-kernel thread
constexpr int kRecSlots = 2 * kMaxJobBands;  // complex per band per record
constexpr size_t kJobCgBytes = 16 * kMaxJobBands;  // per pass and G' row
constexpr int kFreqSlots = 8;     // ach, corb, sch2 (grid part), cor (grid
                                  // part): complex each
constexpr int kRowDoubles = 16;   // per (n1, f) grid-term constants: a
constexpr int kRowHalf = 8;       // pole row and an interval row of 8 each

// One pass over one frequency slice of I_epsR or I_epsA for up to 8 bands:
//   T[b] = sum_{G'} cG[band_b][G'] sum_G aqsn[band_b][G] * slice[G'][G]
// with cG[n1][G'] = vcoul[igp] * conj(aqsm[n1][igp]) (0 for skipped rows).
// Scratch per pass: NG x 8 coefficients (kJobCgBytes per row).
struct SliceJob {
  const double2 *slice;  // [NG][NC]
  int nb;
  int band[kMaxJobBands];
  int pad;
};

// Scalars of the frequency-grid kernel.
struct FreqArgs {
  const double *ekq;        // [NB]
  const double *grid;       // [nFreq]
  const double *pref;       // [nFreq]
  const double2 *brd;       // [nFreq]
  const double2 *schdt;     // [NB][nFreq]
  int NB, NV, nFreq, NE;
  double fmin, fstep, pref_zb;
};

// cGj[(j*NG + G')*8 + b] = vcoul[igp] conj(aqsm[band_b][igp]) for the
// passes jobs[0, njobs) (0 for skipped rows and unused slots).
hipError_t launch_cg(const SliceJob *jobs, int njobs, const double2 *aqsm,
                     const double *vcoul, const int *inv_igp_index,
                     const int *indinv, int NC, int NG, double2 *cGj,
                     hipStream_t s);
// Grid: jobs x row groups x G chunks; rec[(job*ngroups + g)*nchunk + c][16];
// cGj as launch_cg wrote it for the same jobs.
hipError_t launch_slices(const SliceJob *jobs, int njobs, const double2 *cGj,
                         const double2 *aqsn, int NC, int NG, int ngroups,
                         double *rec, hipStream_t s);
// T[job][16] = fixed-order sum of the job's ngroups*nchunk records.
hipError_t launch_slice_reduce(const double *rec, int njobs, int per_job,
                               double *T, hipStream_t s);
int slice_chunks(int NC);
// Number of frequency-grid ranges the grid kernel splits the f sum into.
int freq_splits(int NB, int NE, int nFreq);
// part[((split*NB + n1)*NE + iw)*8]; sem 0 = ffKernel.cpp, 1 = ffKernel.f90;
// rows is scratch of NB*nFreq*kRowDoubles doubles.
hipError_t launch_freq(const FreqArgs &a, int sem, int splits, double *rows,
                       double *part, hipStream_t s);
// out[iw*8] = sum over the nrow = splits*NB rows of part, in order.
hipError_t launch_freq_reduce(const double *part, int nrow, int NE,
                              double *out, hipStream_t s);

// Synthetic I_eps slices on the device (config 1: constants; config 2: the
// counter-based values of ff_fill_inputs).
hipError_t launch_fill_eps(int config, uint64_t seed, int nFreq, int NG,
                           int NC, double2 *epsR, double2 *epsA,
                           hipStream_t s);
hipError_t launch_fill_aqs(int config, uint64_t seed, int NB, int NC,
                           double2 *aqsm, double2 *aqsn, hipStream_t s);

}  // namespace ff
}  // namespace gpp
