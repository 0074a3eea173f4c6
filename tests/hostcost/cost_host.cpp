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
    b.sp0 = cy; b.sp1 = cu; b.sp2 = cv;
    b.rp0 = ry; b.rp1 = ru; b.rp2 = rv;
    b.ss0 = cys; b.ss1 = cus; b.ss2 = cvs;
    b.rs0 = rys; b.rs1 = rus; b.rs2 = rvs;
    b.sl0 = cyl; b.sl1 = cul; b.sl2 = cvl;
    b.rl0 = ryl; b.rl1 = rul; b.rl2 = rvl;
    b.level = level; b.lw = lw; b.lh = lh; b.y_off = y_off; b.c_off = c_off;
    b.clw = lw / 2; b.clh = lh / 2; b.pel = pel; b.satd = satd != 0;
    b.px = px; b.py = py; b.bw = bw; b.bh = bh;
    cand_cost_full(b, mvx, mvy, chroma != 0, *L, *C);
    return 0;
}
