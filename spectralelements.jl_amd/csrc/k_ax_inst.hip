// Instantiates the fused apply kernels for one P (-DSEM_P=<P>) and one part
// (-DSEM_PART): 0 = FP64, 1 = FP32, and for P = 16 2..6 = the FP32-field
// MFMA contraction variants 1..5, so the instantiations compile in parallel.
#include "k_ax_impl.h"

#if !defined(SEM_P) || !defined(SEM_PART)
#error "compile with -DSEM_P=<points per direction> -DSEM_PART=<part>"
#endif

namespace sem {
namespace ax {
#if SEM_PART == 0
template int launch_apply_p<double, SEM_P>(hipStream_t, const Layout&, const ApplyArgs&,
                                           const PcgLaunch*);
#elif SEM_PART == 1
#if SEM_P == 16
extern template void launch_k1<float, 16, 16, 1>(hipStream_t, dim3, dim3, const AxP<float>&, int);
extern template void launch_k1<float, 16, 16, 2>(hipStream_t, dim3, dim3, const AxP<float>&, int);
extern template void launch_k1<float, 16, 16, 3>(hipStream_t, dim3, dim3, const AxP<float>&, int);
extern template void launch_k1<float, 16, 16, 4>(hipStream_t, dim3, dim3, const AxP<float>&, int);
extern template void launch_k1<float, 16, 16, 5>(hipStream_t, dim3, dim3, const AxP<float>&, int);
#endif
template int launch_apply_p<float, SEM_P>(hipStream_t, const Layout&, const ApplyArgs&,
                                          const PcgLaunch*);
#else
template void launch_k1<float, 16, 16, SEM_PART - 1>(hipStream_t, dim3, dim3, const AxP<float>&,
                                                     int);
#endif
}  // namespace ax
}  // namespace sem
