// Host build of the HIP search's whole-candidate cost (open-svpflow_amd/csrc/
// analyse_cost.h) for CPU checks against the oracle's candidate_cost.
#include <cstring>

#include "../../open-svpflow_amd/csrc/analyse_cost.h"

using namespace svpf::cost;

extern "C" int svpf_host_cand_cost(int level, int lw, int lh, int y_off, int c_off, int pel, int satd,
                                   const uint8_t* cy, int cys, int cyl, const uint8_t* cu, int cus, int cul,
                                   const uint8_t* cv, int cvs, int cvl, const uint8_t* ry, int rys, int ryl,
                                   const uint8_t* ru, int rus, int rul, const uint8_t* rv, int rvs, int rvl, int px,
                                   int py, int bw, int bh, int mvx, int mvy, int chroma, uint32_t* L, uint32_t* C) {
    Blk b;
    std::memset(&b, 0, sizeof b);
    b.sp[0] = cy; b.sp[1] = cu; b.sp[2] = cv;
    b.rp[0] = ry; b.rp[1] = ru; b.rp[2] = rv;
    b.ss[0] = cys; b.ss[1] = cus; b.ss[2] = cvs;
    b.rs[0] = rys; b.rs[1] = rus; b.rs[2] = rvs;
    b.sl[0] = cyl; b.sl[1] = cul; b.sl[2] = cvl;
    b.rl[0] = ryl; b.rl[1] = rul; b.rl[2] = rvl;
    b.level = level; b.lw = lw; b.lh = lh; b.y_off = y_off; b.c_off = c_off;
    b.clw = lw / 2; b.clh = lh / 2; b.pel = pel; b.satd = satd != 0;
    b.px = px; b.py = py; b.bw = bw; b.bh = bh;
    cand_cost_full(b, mvx, mvy, chroma != 0, *L, *C);
    return 0;
}
