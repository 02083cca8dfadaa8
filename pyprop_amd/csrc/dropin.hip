// Drop-in replacements for the generated Fortran kernels that PyProp links on the
// basis-propagation path, with their exact C signatures (generator.py:1428-1451: every
// argument by pointer, each array followed by its extents, extent 0 the slowest rank; the
// pair arguments of a Simp rank are int32 [pairCount][2] (row, col)).  Instances:
//   tensorpotentialmultiply_simp_simp_bandnh_   helium (coupled channel x banded-packed x
//                                               banded-nonhermitian), the storage of
//                                               examples/tensor/helium_stabilization
//   tensorpotentialmultiply_diag_bandnh_        hydrogen H0 (reduced spherical diagonal)
//   tensorpotentialmultiply_simp_bandnh_        hydrogen dipole (dipoleselectionrule)
// Semantics as the generated loops: dest += scaling * V * source, dest accumulated.
//
// The reference passes host (blitz) arrays.  Each call stages potential, source and dest
// in device buffers that persist between calls (grown on demand, one set per device and
// host thread), runs pp_tensor_potential_multiply and copies dest back.  This PCIe-bound
// path exists for unmodified callers; pp_tp_create / pp_tp_multiply keep the potential
// resident in HBM.  The generated kernels return nothing and check nothing; a failure here
// (no device, allocation failure) is reported on stderr and leaves dest unchanged.
#include <hip/hip_runtime.h>

#include <complex>
#include <cstdio>
#include <vector>

#include "pp_common.h"
#include "pyprop_amd_dropin.h"

namespace {

struct Staging {
  int device = -1;
  pp::DeviceBuffer<double2> V, src, dst;
};

// never destroyed: the buffers are released by the process's end, not by a destructor that
// could run after the HIP runtime has shut down
thread_local Staging* g_stage_ptr = nullptr;

void multiply(int rank, const int* ids, const std::complex<double>* potential, const int* pext, double scaling,
              const std::complex<double>* source, const int* sext, std::complex<double>* dest, const int* dext,
              const int32_t* const* pairs, const int* counts, const char* name) {
  const int rc = pp::guarded([&] {
    if (!g_stage_ptr) g_stage_ptr = new Staging();
    Staging& g_stage = *g_stage_ptr;
    size_t np = 1, ns = 1, nd = 1;
    for (int r = 0; r < rank; ++r) {
      np *= (size_t)pext[r];
      ns *= (size_t)sext[r];
      nd *= (size_t)dext[r];
    }
    int dev = 0;
    PP_HIP(hipGetDevice(&dev));
    if (g_stage.device != dev) {  // buffers belong to the device that was current
      g_stage.V.release();
      g_stage.src.release();
      g_stage.dst.release();
      g_stage.device = dev;
    }
    if (g_stage.V.count < np) g_stage.V.resize(np);
    if (g_stage.src.count < ns) g_stage.src.resize(ns);
    if (g_stage.dst.count < nd) g_stage.dst.resize(nd);
    PP_HIP(hipMemcpy(g_stage.V.ptr, potential, np * sizeof(double2), hipMemcpyHostToDevice));
    PP_HIP(hipMemcpy(g_stage.src.ptr, source, ns * sizeof(double2), hipMemcpyHostToDevice));
    PP_HIP(hipMemcpy(g_stage.dst.ptr, dest, nd * sizeof(double2), hipMemcpyHostToDevice));
    if (pp_tensor_potential_multiply(rank, ids, reinterpret_cast<const pp_cplx*>(g_stage.V.ptr), pext, scaling,
                                     reinterpret_cast<const pp_cplx*>(g_stage.src.ptr), sext,
                                     reinterpret_cast<pp_cplx*>(g_stage.dst.ptr), dext, pairs, counts,
                                     nullptr) != 0)
      throw pp::Error(PP_ERR_INVALID, pp_last_error());
    PP_HIP(hipMemcpy(dest, g_stage.dst.ptr, nd * sizeof(double2), hipMemcpyDeviceToHost));
  });
  if (rc != 0) std::fprintf(stderr, "pyprop_amd: %s failed: %s\n", name, pp_last_error());
}

}  // namespace

extern "C" void tensorpotentialmultiply_simp_simp_bandnh_(
    std::complex<double>* potential, int* potentialExtent0, int* potentialExtent1, int* potentialExtent2,
    double* scaling, std::complex<double>* source, int* sourceExtent0, int* sourceExtent1, int* sourceExtent2,
    std::complex<double>* dest, int* destExtent0, int* destExtent1, int* destExtent2, int* pair0, int* pair0Extent0,
    int* pair0Extent1, int* pair1, int* pair1Extent0, int* pair1Extent1) {
  const int ids[3] = {PP_STORAGE_SIMP, PP_STORAGE_SIMP, PP_STORAGE_BANDNH};
  const int pext[3] = {*potentialExtent0, *potentialExtent1, *potentialExtent2};
  const int sext[3] = {*sourceExtent0, *sourceExtent1, *sourceExtent2};
  const int dext[3] = {*destExtent0, *destExtent1, *destExtent2};
  if (*pair0Extent1 != 2 || *pair1Extent1 != 2) {
    std::fprintf(stderr, "pyprop_amd: tensorpotentialmultiply_simp_simp_bandnh_: pair arrays must be [count][2]\n");
    return;
  }
  const int32_t* pairs[3] = {pair0, pair1, nullptr};
  const int counts[3] = {*pair0Extent0, *pair1Extent0, 0};
  multiply(3, ids, potential, pext, *scaling, source, sext, dest, dext, pairs, counts,
           "tensorpotentialmultiply_simp_simp_bandnh_");
}

extern "C" void tensorpotentialmultiply_diag_bandnh_(std::complex<double>* potential, int* potentialExtent0,
                                                     int* potentialExtent1, double* scaling,
                                                     std::complex<double>* source, int* sourceExtent0,
                                                     int* sourceExtent1, std::complex<double>* dest,
                                                     int* destExtent0, int* destExtent1) {
  const int ids[2] = {PP_STORAGE_DIAG, PP_STORAGE_BANDNH};
  const int pext[2] = {*potentialExtent0, *potentialExtent1};
  const int sext[2] = {*sourceExtent0, *sourceExtent1};
  const int dext[2] = {*destExtent0, *destExtent1};
  const int32_t* pairs[2] = {nullptr, nullptr};
  const int counts[2] = {0, 0};
  multiply(2, ids, potential, pext, *scaling, source, sext, dest, dext, pairs, counts,
           "tensorpotentialmultiply_diag_bandnh_");
}

extern "C" void tensorpotentialmultiply_simp_bandnh_(std::complex<double>* potential, int* potentialExtent0,
                                                     int* potentialExtent1, double* scaling,
                                                     std::complex<double>* source, int* sourceExtent0,
                                                     int* sourceExtent1, std::complex<double>* dest,
                                                     int* destExtent0, int* destExtent1, int* pair0,
                                                     int* pair0Extent0, int* pair0Extent1) {
  const int ids[2] = {PP_STORAGE_SIMP, PP_STORAGE_BANDNH};
  const int pext[2] = {*potentialExtent0, *potentialExtent1};
  const int sext[2] = {*sourceExtent0, *sourceExtent1};
  const int dext[2] = {*destExtent0, *destExtent1};
  if (*pair0Extent1 != 2) {
    std::fprintf(stderr, "pyprop_amd: tensorpotentialmultiply_simp_bandnh_: pair array must be [count][2]\n");
    return;
  }
  const int32_t* pairs[2] = {pair0, nullptr};
  const int counts[2] = {*pair0Extent0, 0};
  multiply(2, ids, potential, pext, *scaling, source, sext, dest, dext, pairs, counts,
           "tensorpotentialmultiply_simp_bandnh_");
}
