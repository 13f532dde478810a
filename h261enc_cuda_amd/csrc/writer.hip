/*
 * writer.hip -- the H.261 picture writer on the host.
 *
 * Same syntax and output bytes as the reference writer
 * (src/h261_picture.c:38-71, src/h261_gob.c:6-199,
 * src/h261_macroblock.c:5-262, src/h261_bitstream.c:28-77): picture
 * header, 12 (CIF) or 3 (QCIF) GOBs, MBA/MTYPE/MQUANT/MVD/CBP and TCOEFF
 * with the escape code.  The reference zeroes a 1 MiB buffer per frame,
 * writes one bit at a time and searches the run/level table linearly; this
 * writer packs codes into a 64-bit accumulator and indexes the table by
 * (run, level).
 */
#include <stdint.h>
#include <string.h>

#include "../../include/h261gpu.h"
#include "h261_vlc.h"

namespace h261g {

namespace {

const Vlc k_mba[35] = H261G_MBA_TABLE;
const Vlc k_mtype[10] = H261G_MTYPE_TABLE;
const Vlc k_cbp[64] = H261G_CBP_TABLE;
const Vlc k_mvd[32] = H261G_MVD_TABLE;
const Vlc k_tcoeff[kTcoeffRuns][kTcoeffLevels] = H261G_TCOEFF_TABLE;
const uint8_t k_zigzag[64] = H261G_ZIGZAG;
const int k_cbp_bit[6] = {32, 16, 8, 4, 2, 1};

struct BitWriter {
    uint8_t *out;
    int64_t cap;
    int64_t pos = 0; /* bytes flushed */
    uint64_t acc = 0;
    int nacc = 0;
    bool overflow = false;
    bool invalid = false;

    void put(uint32_t code, int len)
    {
        if (len <= 0) return;
        code &= len >= 32 ? 0xffffffffu : ((1u << len) - 1u);
        acc = (acc << len) | code;
        nacc += len;
        while (nacc >= 8) {
            nacc -= 8;
            if (pos < cap)
                out[pos] = (uint8_t)(acc >> nacc);
            else
                overflow = true;
            pos++;
        }
    }
    int64_t finish()
    {
        if (nacc > 0) {
            if (pos < cap)
                out[pos] = (uint8_t)(acc << (8 - nacc));
            else
                overflow = true;
            pos++;
            nacc = 0;
        }
        return pos;
    }
};

/* src/h261_macroblock.c:5-92 */
void put_tcoeff(BitWriter &w, const Int32 *c, bool intra)
{
    int run, start;
    if (intra) {
        w.put((uint32_t)c[0], 8);
        run = 0;
        start = 1;
    } else if (c[0]) {
        if (c[0] == 1 || c[0] == -1) {
            w.put(1, 1);
            w.put(c[0] < 0, 1);
            start = 1;
        } else
            start = 0;
        run = 0;
    } else {
        run = 1;
        start = 1;
    }
    for (int i = start; i < 64; i++) {
        const int l = c[k_zigzag[i]];
        if (!l) {
            run++;
            continue;
        }
        const int a = l < 0 ? -l : l;
        const Vlc *v = (run < kTcoeffRuns && a < kTcoeffLevels) ? &k_tcoeff[run][a] : nullptr;
        if (v && v->len) {
            w.put(v->code, v->len);
            w.put(l < 0, 1);
        } else {
            w.put(1, 6);
            w.put((uint32_t)run, 6);
            w.put((uint32_t)l, 8);
        }
        run = 0;
    }
    if (run == 64) w.invalid = true; /* a coded block with no coefficient */
    w.put(2, 2);                     /* EOB */
}

struct CodingVars {
    int quant, mv_valid, mv_x, mv_y;
};

/* src/h261_macroblock.c:94-262 */
void put_macroblock(BitWriter &w, const h261_macroblock_t &mb, int addr, const CodingVars &cv)
{
    if (mb.i_macroblock_type == H261_MB_TYPE_INTRA) {
        w.put(k_mba[addr].code, k_mba[addr].len);
        if (mb.i_macroblock_quant != cv.quant) {
            w.put(k_mtype[1].code, k_mtype[1].len);
            w.put((uint32_t)mb.i_macroblock_quant, 5);
        } else
            w.put(k_mtype[0].code, k_mtype[0].len);
        for (int b = 0; b < 6; b++) put_tcoeff(w, mb.rgi_tcoeff[b], true);
        return;
    }
    if (mb.i_macroblock_type != H261_MB_TYPE_INTER) return;
    w.put(k_mba[addr].code, k_mba[addr].len);
    static const int t0[8] = {-1, 4, -1, 7, 2, 5, -1, 8}, t1[8] = {-1, -1, -1, -1, 3, 6, -1, 9};
    const int flags = mb.i_macroblock_type_flags;
    const bool cq = mb.i_macroblock_quant != cv.quant && (flags & H261_MB_COEFF);
    const int t = (flags >= 0 && flags < 8) ? (cq ? t1[flags] : t0[flags]) : -1;
    if (t < 0) {
        w.invalid = true;
        return;
    }
    w.put(k_mtype[t].code, k_mtype[t].len);
    if (cq) w.put((uint32_t)mb.i_macroblock_quant, 5);
    if (flags & H261_MB_MC) {
        int dx = mb.rgi_mv[0], dy = mb.rgi_mv[1];
        if (cv.mv_valid) {
            dx -= cv.mv_x;
            dy -= cv.mv_y;
        }
        if (dx < -30 || dx > 30 || dy < -30 || dy > 30) {
            w.invalid = true;
            return;
        }
        const Vlc &vx = k_mvd[(dx + 48) & 31], &vy = k_mvd[(dy + 48) & 31];
        w.put(vx.code, vx.len);
        w.put(vy.code, vy.len);
    }
    if (flags & H261_MB_COEFF) {
        const int cbp = mb.i_coded_block_pattern;
        if (cbp <= 0 || cbp > 63) {
            w.invalid = true;
            return;
        }
        w.put(k_cbp[cbp].code, k_cbp[cbp].len);
        for (int b = 0; b < 6; b++)
            if (cbp & k_cbp_bit[b]) put_tcoeff(w, mb.rgi_tcoeff[b], false);
    }
}

}  // namespace

}  // namespace h261g

extern "C" int h261g_write_picture(const h261_macroblock_t *mbs, int32_t width, int32_t height, int32_t frame_num,
                                   int32_t quant, uint8_t *out, int32_t cap)
{
    using namespace h261g;
    const bool cif = width == 352 && height == 288, qcif = width == 176 && height == 144;
    if (!mbs || !out || cap < 0 || !(cif || qcif) || quant < 1 || quant > 31) return H261G_ERR_ARG;
    const int mb_w = width / 16;
    BitWriter w;
    w.out = out;
    w.cap = cap;
    /* picture header (picture.c:38-71, field values from picture.c:5-35) */
    w.put(16, 20);
    w.put((uint32_t)(frame_num & 31), 5);
    w.put(0, 1); /* split screen */
    w.put(0, 1); /* document camera */
    w.put(1, 1); /* freeze picture release */
    w.put(cif ? 1 : 0, 1);
    w.put(1, 1); /* hi-res bit as the reference sets it */
    w.put(1, 1); /* spare */
    w.put(0, 1); /* PEI */
    const int n_gob = cif ? 12 : 3;
    for (int g = 0; g < n_gob; g++) {
        const int gn = cif ? g + 1 : 2 * g + 1;
        const int gx = cif ? (g & 1) * 11 : 0, gy = cif ? (g >> 1) * 3 : g * 3;
        w.put(1, 16);
        w.put((uint32_t)gn, 4);
        w.put((uint32_t)quant, 5);
        w.put(0, 1);
        CodingVars cv = {quant, 0, 0, 0};
        int addr = 0;
        for (int i = 0; i < 33; i++) {
            addr++;
            if (i == 0 || i == 11 || i == 22) cv.mv_valid = 0;
            const h261_macroblock_t &mb = mbs[(gy + i / 11) * mb_w + gx + i % 11];
            const int flags = mb.i_macroblock_type_flags;
            if (!flags) {
                cv.mv_valid = 0;
                continue;
            }
            put_macroblock(w, mb, addr, cv);
            addr = 0;
            if (flags & H261_MB_COEFF) cv.quant = mb.i_macroblock_quant;
            if (flags & H261_MB_MC) {
                cv.mv_x = mb.rgi_mv[0];
                cv.mv_y = mb.rgi_mv[1];
                cv.mv_valid = 1;
            } else
                cv.mv_valid = 0;
        }
    }
    const int64_t n = w.finish();
    if (w.invalid) return H261G_ERR_ARG;
    if (w.overflow) return H261G_ERR_FULL;
    return (int)n;
}
