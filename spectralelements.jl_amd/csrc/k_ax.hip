// Dispatch of the fused apply over P (2..16); the kernels are instantiated
// per P in k_ax_inst.hip.
#include "k_ax_impl.h"

namespace sem {
namespace ax {
#define SEM_AX_EXTERN(P)                                                                        \
  extern template int launch_apply_p<double, P>(hipStream_t, const Layout&, const ApplyArgs&, \
                                                const PcgLaunch*);                            \
  extern template int launch_apply_p<float, P>(hipStream_t, const Layout&, const ApplyArgs&,  \
                                               const PcgLaunch*);
SEM_AX_EXTERN(2) SEM_AX_EXTERN(3) SEM_AX_EXTERN(4) SEM_AX_EXTERN(5) SEM_AX_EXTERN(6)
SEM_AX_EXTERN(7) SEM_AX_EXTERN(8) SEM_AX_EXTERN(9) SEM_AX_EXTERN(10) SEM_AX_EXTERN(11)
SEM_AX_EXTERN(12) SEM_AX_EXTERN(13) SEM_AX_EXTERN(14) SEM_AX_EXTERN(15) SEM_AX_EXTERN(16)
#undef SEM_AX_EXTERN

template <typename T>
int dispatch(hipStream_t st, const Layout& g, const ApplyArgs& a, const PcgLaunch* pc) {
  switch (g.P) {
    case 2: return launch_apply_p<T, 2>(st, g, a, pc);
    case 3: return launch_apply_p<T, 3>(st, g, a, pc);
    case 4: return launch_apply_p<T, 4>(st, g, a, pc);
    case 5: return launch_apply_p<T, 5>(st, g, a, pc);
    case 6: return launch_apply_p<T, 6>(st, g, a, pc);
    case 7: return launch_apply_p<T, 7>(st, g, a, pc);
    case 8: return launch_apply_p<T, 8>(st, g, a, pc);
    case 9: return launch_apply_p<T, 9>(st, g, a, pc);
    case 10: return launch_apply_p<T, 10>(st, g, a, pc);
    case 11: return launch_apply_p<T, 11>(st, g, a, pc);
    case 12: return launch_apply_p<T, 12>(st, g, a, pc);
    case 13: return launch_apply_p<T, 13>(st, g, a, pc);
    case 14: return launch_apply_p<T, 14>(st, g, a, pc);
    case 15: return launch_apply_p<T, 15>(st, g, a, pc);
    case 16: return launch_apply_p<T, 16>(st, g, a, pc);
    default:
      set_error("fused apply supports 2 <= nr == ns <= 16 (got %d)", g.P);
      return SEM_ERR_UNSUPPORTED;
  }
}

}  // namespace ax

AxPlan make_plan(const Layout& g, bool assemble, bool pcg, int rows) {
  AxPlan pl{};
  const int cx = SEM_AX_THREADS / g.P;
  pl.S = assemble ? ax::rows_per_segment(g, cx, pcg, rows) : 1;
  // The assembled FMA apply (P < 16; 16 runs the MFMA forms) takes two
  // element rows side by side where one-row workgroups would be used, the
  // PCG iteration where a two-row sweep would be (it also runs three waves
  // per SIMD there instead of two). rows < 0 forces RY = -rows (test hook:
  // 2 or 4; 2 in PCG mode).
  pl.RY = 1;
  if (assemble && g.P < 16) {
    if (rows < 0) {
      const int ry = -rows;
      if ((ry == 2 || (ry == 4 && !pcg)) && cx % ry == 0) pl.RY = ry;
    } else if (rows == 0 && pl.S == (pcg ? 2 : 1) && cx % 2 == 0 && g.Ey >= 2) {
      pl.RY = 2;
    }
  }
  if (pl.RY > 1) pl.S = pl.RY;
  pl.CX = cx / pl.RY;
  pl.nstrips = (g.Ex + pl.CX - 1) / pl.CX;
  pl.nsegs = (g.Ey + pl.S - 1) / pl.S;
  pl.nk1 = pl.nstrips * pl.nsegs;
  if (assemble) {
    const int nxs = pl.nstrips - 1 + (g.per_x ? 1 : 0);
    const int nys = pl.nsegs - 1 + (g.per_y ? 1 : 0);
    pl.nxitems = (long)nxs * g.NC;
    pl.total = pl.nxitems + (long)nys * g.L;
    pl.nk2 = (int)((pl.total + 255) / 256);
  }
  return pl;
}

template <typename T>
int launch_apply(hipStream_t st, const Layout& g, const ApplyArgs& a) {
  return ax::dispatch<T>(st, g, a, nullptr);
}

template <typename T>
int launch_apply_pcg(hipStream_t st, const Layout& g, const ApplyArgs& a, const PcgLaunch& pc) {
  return ax::dispatch<T>(st, g, a, &pc);
}

template int launch_apply<double>(hipStream_t, const Layout&, const ApplyArgs&);
template int launch_apply<float>(hipStream_t, const Layout&, const ApplyArgs&);
template int launch_apply_pcg<double>(hipStream_t, const Layout&, const ApplyArgs&,
                                      const PcgLaunch&);
template int launch_apply_pcg<float>(hipStream_t, const Layout&, const ApplyArgs&,
                                     const PcgLaunch&);

}  // namespace sem
