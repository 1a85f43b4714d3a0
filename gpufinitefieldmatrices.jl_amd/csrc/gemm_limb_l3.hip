// gemm_limb_l3.hip -- limb-form GEMM launchers for 3 digits per residue.
#include "gemm_limb.h"

template int launch_limb_impl<3>(modp_ctx *ctx, const GemmArgs &g, int epi, const int8_t *Al, int64_t Mp,
                                  const int8_t *Bl, int64_t Np, int64_t Kp, const LimbW &lw, int tile_m,
                                  int tile_n, const Slices &sl);
