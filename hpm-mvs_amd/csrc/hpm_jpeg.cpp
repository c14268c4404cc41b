/*
 * hpm_jpeg.cpp — the grey image of a JPEG file, as the reference reads its
 * images: cv::imread(path, cv::IMREAD_GRAYSCALE) (HPM-MVS/HPM.cpp:492-494),
 * which asks libjpeg for grayscale output, i.e. the decoded luminance (Y)
 * component, computed with libjpeg's default integer inverse DCT.
 *
 * Supported: baseline, extended sequential and progressive Huffman JPEG
 * (SOF0, SOF1, SOF2), 8-bit samples, 1 (grey) or 3 (YCbCr) components with
 * any sampling factors, interleaved or single-component scans, restart
 * intervals. Progressive scans follow ITU-T T.81 annex G: DC first and
 * refinement scans, AC spectral-selection first scans with end-of-band
 * runs, and AC successive-approximation refinement (correction bits of
 * the coefficients already nonzero, new coefficients of +-1 at the current
 * bit position). For the grey image only the Y component is reconstructed
 * (it is at full resolution in every sampling layout). Lossless,
 * arithmetic-coded, 12-bit and 4-component (CMYK/YCCK) files are
 * rejected.
 *
 * The inverse DCT is the integer form of the Loeffler-Ligtenberg-Moschytz
 * factorisation (ICASSP 1989) with 13-bit fixed-point constants and two
 * extra bits of precision between the column and the row pass, rounded
 * right shifts, and the output level-shifted by 128 and clamped to 0..255,
 * as libjpeg's "islow" method computes it.
 *
 * The colour image (cv::imread(path, cv::IMREAD_COLOR), read by RunFusion
 * for the point colours, HPM-MVS/main.cpp:523) reconstructs every
 * component, upsamples 2x1 / 2x2 subsampled chroma with libjpeg's default
 * triangle filter and converts YCbCr to RGB in libjpeg's fixed point.
 */
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "hpm_io.h"

namespace {

/* zig-zag position k -> natural (row-major) index */
const uint8_t kZigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                             12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                             35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                             58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

struct Huffman {
    /* canonical code tables: for each length L, the first code, the index of
     * its first symbol and the number of codes */
    int mincode[17], maxcode[18], valptr[17];
    uint8_t vals[256];
    bool defined = false;
};

struct Component {
    int id = 0, h = 1, v = 1, tq = 0;
    int td = 0, ta = 0;
    int bw = 0, bh = 0; /* blocks per row / column of the component */
    int pred = 0;       /* DC predictor */
    std::vector<int16_t> coef; /* bw * bh * 64, natural order, quantised */
};

struct Decoder {
    const uint8_t* p = nullptr;
    const uint8_t* end = nullptr;
    uint16_t q[4][64];
    bool qdef[4] = {false, false, false, false};
    Huffman dc[4], ac[4];
    std::vector<Component> comps;
    int width = 0, height = 0, hmax = 1, vmax = 1;
    int restart = 0;
    bool progressive = false;
    int eobrun = 0; /* progressive AC scans: blocks left in an end-of-band run */
    /* bit reader */
    uint32_t bits = 0;
    int nbits = 0;
    bool hit_marker = false;
    std::string err;

    bool fail(const char* m)
    {
        if (err.empty()) {
            err = m;
        }
        return false;
    }

    int byte()
    {
        return p < end ? *p++ : -1;
    }

    int u16()
    {
        const int a = byte(), b = byte();
        return (a < 0 || b < 0) ? -1 : (a << 8) | b;
    }

    /* entropy-coded data: FF 00 is a data FF; any other marker ends the
     * data (zeros are fed after it) */
    void fill()
    {
        while (nbits <= 24) {
            int c = 0;
            if (!hit_marker && p < end) {
                c = *p;
                if (c == 0xFF) {
                    const int n = p + 1 < end ? p[1] : -1;
                    if (n == 0x00) {
                        p += 2;
                    } else {
                        hit_marker = true;
                        c = 0;
                    }
                } else {
                    ++p;
                }
            }
            bits |= (uint32_t)c << (24 - nbits);
            nbits += 8;
        }
    }

    int getbit()
    {
        if (nbits == 0) {
            fill();
        }
        const int b = (int)(bits >> 31);
        bits <<= 1;
        --nbits;
        return b;
    }

    int getbits(int n)
    {
        if (n == 0) {
            return 0;
        }
        if (nbits < n) {
            fill();
        }
        const int v = (int)(bits >> (32 - n));
        bits <<= n;
        nbits -= n;
        return v;
    }

    /* value of an n-bit magnitude category (the EXTEND procedure of ITU-T
     * T.81 F.2.2.1) */
    static int extend(int v, int n)
    {
        return v < (1 << (n - 1)) ? v - (1 << n) + 1 : v;
    }

    int decode(const Huffman& t)
    {
        int code = 0;
        for (int l = 1; l <= 16; ++l) {
            code = (code << 1) | getbit();
            if (code <= t.maxcode[l]) {
                return t.vals[t.valptr[l] + code - t.mincode[l]];
            }
        }
        return -1; /* corrupt */
    }

    bool read_dqt(int len)
    {
        const uint8_t* stop = p + len;
        while (p < stop) {
            const int pq_tq = byte();
            const int pq = pq_tq >> 4, tq = pq_tq & 15;
            if (tq > 3 || pq > 1) {
                return fail("bad quantisation table");
            }
            for (int k = 0; k < 64; ++k) {
                const int v = pq ? u16() : byte();
                if (v < 0) {
                    return fail("truncated quantisation table");
                }
                q[tq][kZigzag[k]] = (uint16_t)v;
            }
            qdef[tq] = true;
        }
        return p == stop || fail("bad DQT length");
    }

    bool read_dht(int len)
    {
        const uint8_t* stop = p + len;
        while (p < stop) {
            const int tc_th = byte();
            const int tc = tc_th >> 4, th = tc_th & 15;
            if (tc > 1 || th > 3) {
                return fail("bad Huffman table");
            }
            Huffman& t = tc ? ac[th] : dc[th];
            int counts[17] = {0};
            int total = 0;
            for (int l = 1; l <= 16; ++l) {
                counts[l] = byte();
                total += counts[l];
            }
            if (total > 256 || p + total > stop) {
                return fail("bad Huffman table counts");
            }
            for (int i = 0; i < total; ++i) {
                t.vals[i] = (uint8_t)byte();
            }
            /* canonical codes (ITU-T T.81 C.2) */
            int code = 0, k = 0;
            for (int l = 1; l <= 16; ++l) {
                t.valptr[l] = k;
                t.mincode[l] = code;
                code += counts[l];
                k += counts[l];
                t.maxcode[l] = counts[l] ? code - 1 : -1;
                code <<= 1;
            }
            t.maxcode[17] = 0x7fffffff;
            t.defined = true;
        }
        return p == stop || fail("bad DHT length");
    }

    bool read_sof(int len, int marker)
    {
        if (marker != 0xC0 && marker != 0xC1 && marker != 0xC2) {
            return fail("unsupported JPEG process (only baseline, extended sequential and progressive Huffman)");
        }
        progressive = marker == 0xC2;
        const int prec = byte();
        height = u16();
        width = u16();
        const int nc = byte();
        if (prec != 8) {
            return fail("only 8-bit JPEG samples are supported");
        }
        if (height <= 0 || width <= 0) {
            return fail("JPEG size missing (DNL is not supported)");
        }
        if (nc != 1 && nc != 3) {
            return fail("only grey or YCbCr JPEG images are supported");
        }
        if (len != 6 + 3 * nc) {
            return fail("bad SOF length");
        }
        comps.assign((size_t)nc, Component());
        for (int i = 0; i < nc; ++i) {
            Component& c = comps[(size_t)i];
            c.id = byte();
            const int hv = byte();
            c.h = hv >> 4;
            c.v = hv & 15;
            c.tq = byte();
            if (c.h < 1 || c.h > 4 || c.v < 1 || c.v > 4 || c.tq > 3) {
                return fail("bad component in SOF");
            }
            hmax = std::max(hmax, c.h);
            vmax = std::max(vmax, c.v);
        }
        const int mcux = (width + 8 * hmax - 1) / (8 * hmax), mcuy = (height + 8 * vmax - 1) / (8 * vmax);
        for (Component& c : comps) {
            c.bw = mcux * c.h;
            c.bh = mcuy * c.v;
            c.coef.assign((size_t)c.bw * c.bh * 64, 0);
        }
        return true;
    }

    bool decode_block(Component& c, int16_t* blk)
    {
        const Huffman& hd = dc[c.td];
        const Huffman& ha = ac[c.ta];
        const int t = decode(hd);
        if (t < 0 || t > 11) {
            return fail("corrupt DC code");
        }
        const int diff = t ? extend(getbits(t), t) : 0;
        c.pred += diff;
        blk[0] = (int16_t)c.pred;
        for (int k = 1; k < 64;) {
            const int rs = decode(ha);
            if (rs < 0) {
                return fail("corrupt AC code");
            }
            const int r = rs >> 4, s = rs & 15;
            if (s == 0) {
                if (r == 15) { /* ZRL: 16 zeros */
                    k += 16;
                    continue;
                }
                break; /* EOB */
            }
            k += r;
            if (k > 63) {
                return fail("AC run past the block");
            }
            blk[kZigzag[k]] = (int16_t)extend(getbits(s), s);
            ++k;
        }
        return true;
    }

    /* progressive DC first scan: the DC difference, scaled by 2^al */
    bool decode_dc_first(Component& c, int16_t* blk, int al)
    {
        const int t = decode(dc[c.td]);
        if (t < 0 || t > 11) {
            return fail("corrupt DC code");
        }
        c.pred += t ? extend(getbits(t), t) : 0;
        blk[0] = (int16_t)(c.pred * (1 << al));
        return true;
    }

    /* progressive DC refinement: one more bit of the DC coefficient */
    void decode_dc_refine(int16_t* blk, int al)
    {
        if (getbits(1)) {
            blk[0] = (int16_t)(blk[0] | (1 << al));
        }
    }

    /* progressive AC first scan over the band ss..se, values scaled by 2^al,
     * with end-of-band runs across blocks */
    bool decode_ac_first(Component& c, int16_t* blk, int ss, int se, int al)
    {
        if (eobrun > 0) {
            --eobrun;
            return true;
        }
        const Huffman& ha = ac[c.ta];
        for (int k = ss; k <= se; ++k) {
            const int rs = decode(ha);
            if (rs < 0) {
                return fail("corrupt AC code");
            }
            const int r = rs >> 4, sz = rs & 15;
            if (sz == 0) {
                if (r < 15) { /* EOBr: this block and (2^r - 1 + bits) more end here */
                    eobrun = (1 << r) - 1;
                    if (r) {
                        eobrun += getbits(r);
                    }
                    break;
                }
                k += 15; /* ZRL */
                continue;
            }
            k += r;
            if (k > se) {
                return fail("AC run past the band");
            }
            blk[kZigzag[k]] = (int16_t)(extend(getbits(sz), sz) * (1 << al));
        }
        return true;
    }

    /* progressive AC refinement of the band ss..se at bit al: a correction
     * bit for every coefficient already nonzero that the scan passes, and
     * new coefficients of magnitude 2^al after runs of zero coefficients */
    bool decode_ac_refine(Component& c, int16_t* blk, int ss, int se, int al)
    {
        const int p1 = 1 << al, m1 = -(1 << al);
        auto correct = [&](int k) {
            int16_t& v = blk[kZigzag[k]];
            if (getbits(1) && (v & p1) == 0) {
                v = (int16_t)(v >= 0 ? v + p1 : v + m1);
            }
        };
        int k = ss;
        if (eobrun == 0) {
            const Huffman& ha = ac[c.ta];
            for (; k <= se; ++k) {
                const int rs = decode(ha);
                if (rs < 0) {
                    return fail("corrupt AC code");
                }
                int r = rs >> 4;
                const int sz = rs & 15;
                int val = 0;
                if (sz) {
                    if (sz != 1) {
                        return fail("bad AC refinement value");
                    }
                    val = getbits(1) ? p1 : m1;
                } else if (r != 15) { /* EOBr */
                    eobrun = 1 << r;
                    if (r) {
                        eobrun += getbits(r);
                    }
                    break;
                }
                /* pass r zero coefficients (correcting the nonzero ones on
                 * the way), then place the new value (ZRL: 16 zeros) */
                for (; k <= se; ++k) {
                    if (blk[kZigzag[k]] != 0) {
                        correct(k);
                    } else {
                        if (r == 0) {
                            break;
                        }
                        --r;
                    }
                }
                if (val && k <= se) {
                    blk[kZigzag[k]] = (int16_t)val;
                }
            }
        }
        if (eobrun > 0) {
            /* inside an end-of-band run: correction bits only */
            for (; k <= se; ++k) {
                if (blk[kZigzag[k]] != 0) {
                    correct(k);
                }
            }
            --eobrun;
        }
        return true;
    }

    /* after a restart interval: the RSTn marker, then fresh predictors */
    bool restart_marker()
    {
        bits = 0;
        nbits = 0;
        hit_marker = false;
        while (p + 1 < end && !(p[0] == 0xFF && p[1] >= 0xD0 && p[1] <= 0xD7)) {
            ++p; /* tolerate padding before the marker */
        }
        if (p + 1 >= end) {
            return fail("missing restart marker");
        }
        p += 2;
        for (Component& c : comps) {
            c.pred = 0;
        }
        eobrun = 0;
        return true;
    }

    bool read_sos(int len)
    {
        const int ns = byte();
        if (ns < 1 || ns > 4 || len != 4 + 2 * ns) {
            return fail("bad SOS");
        }
        std::vector<Component*> sc;
        for (int i = 0; i < ns; ++i) {
            const int id = byte(), t = byte();
            Component* c = nullptr;
            for (Component& k : comps) {
                if (k.id == id) {
                    c = &k;
                }
            }
            if (!c) {
                return fail("scan names an unknown component");
            }
            c->td = t >> 4;
            c->ta = t & 15;
            if (c->td > 3 || c->ta > 3) {
                return fail("scan uses an undefined Huffman table");
            }
            sc.push_back(c);
        }
        const int ss = byte(), se = byte(), ahal = byte();
        const int ah = ahal >> 4, al = ahal & 15;
        if (!progressive) {
            if (ss != 0 || se != 63 || ahal != 0) {
                return fail("not a sequential scan");
            }
        } else if (ss > se || se > 63 || al > 13 || (ss == 0 && se != 0) || (ss > 0 && ns != 1)) {
            return fail("bad progressive scan");
        }
        /* the Huffman tables this scan decodes with must be defined */
        const bool dc_scan = ss == 0, first = ah == 0;
        for (const Component* c : sc) {
            if ((dc_scan && first && !dc[c->td].defined) || (!dc_scan && !ac[c->ta].defined) ||
                (!progressive && (!dc[c->td].defined || !ac[c->ta].defined))) {
                return fail("scan uses an undefined Huffman table");
            }
        }
        eobrun = 0;
        auto block = [&](Component& c, int16_t* blk) {
            if (!progressive) {
                return decode_block(c, blk);
            }
            if (dc_scan) {
                if (first) {
                    return decode_dc_first(c, blk, al);
                }
                decode_dc_refine(blk, al);
                return true;
            }
            return first ? decode_ac_first(c, blk, ss, se, al) : decode_ac_refine(c, blk, ss, se, al);
        };
        bits = 0;
        nbits = 0;
        hit_marker = false;
        for (Component* c : sc) {
            c->pred = 0;
        }
        int mcus_x, mcus_y;
        if (ns == 1) { /* non-interleaved: one block per MCU over the component's own extent */
            Component& c = *sc[0];
            const int cw = (width * c.h + hmax - 1) / hmax, ch = (height * c.v + vmax - 1) / vmax;
            mcus_x = (cw + 7) / 8;
            mcus_y = (ch + 7) / 8;
        } else {
            mcus_x = (width + 8 * hmax - 1) / (8 * hmax);
            mcus_y = (height + 8 * vmax - 1) / (8 * vmax);
        }
        int todo = restart;
        for (int my = 0; my < mcus_y; ++my) {
            for (int mx = 0; mx < mcus_x; ++mx) {
                if (restart && todo == 0) {
                    if (!restart_marker()) {
                        return false;
                    }
                    todo = restart;
                }
                if (ns == 1) {
                    Component& c = *sc[0];
                    if (!block(c, &c.coef[((size_t)my * c.bw + mx) * 64])) {
                        return false;
                    }
                } else {
                    for (Component* c : sc) {
                        for (int by = 0; by < c->v; ++by) {
                            for (int bx = 0; bx < c->h; ++bx) {
                                const size_t b = (size_t)(my * c->v + by) * c->bw + (size_t)(mx * c->h + bx);
                                if (!block(*c, &c->coef[b * 64])) {
                                    return false;
                                }
                            }
                        }
                    }
                }
                --todo;
            }
        }
        /* leave p at the marker that ended the entropy-coded data */
        while (p + 1 < end && !(p[0] == 0xFF && p[1] != 0x00 && !(p[1] >= 0xD0 && p[1] <= 0xD7))) {
            ++p;
        }
        return true;
    }

    bool parse()
    {
        if (u16() != 0xFFD8) {
            return fail("not a JPEG file");
        }
        bool frame = false, scanned = false;
        while (p < end) {
            int c = byte();
            if (c != 0xFF) {
                continue; /* garbage between segments */
            }
            int m = byte();
            while (m == 0xFF) {
                m = byte();
            }
            if (m < 0) {
                break;
            }
            if (m == 0xD9) { /* EOI */
                break;
            }
            if (m == 0x01 || (m >= 0xD0 && m <= 0xD7)) {
                continue;
            }
            const int len = u16();
            if (len < 2 || p + len - 2 > end) {
                return fail("truncated JPEG segment");
            }
            const int body = len - 2;
            if (m == 0xDB) {
                if (!read_dqt(body)) return false;
            } else if (m == 0xC4) {
                if (!read_dht(body)) return false;
            } else if (m == 0xDD) {
                restart = u16();
            } else if ((m >= 0xC0 && m <= 0xCF) && m != 0xC4 && m != 0xC8 && m != 0xCC) {
                if (frame) return fail("more than one frame");
                if (!read_sof(body, m)) return false;
                frame = true;
            } else if (m == 0xCC) {
                return fail("arithmetic coding is not supported");
            } else if (m == 0xDA) {
                if (!frame) return fail("scan before frame");
                if (!read_sos(body)) return false;
                scanned = true;
            } else {
                p += body; /* APPn, COM, DNL, ... */
            }
        }
        if (!scanned) {
            return fail("no scan in JPEG file");
        }
        for (const Component& c : comps) {
            if (!qdef[c.tq]) {
                return fail("component uses an undefined quantisation table");
            }
        }
        return true;
    }
};

/* integer inverse DCT of one dequantised block into 8x8 samples */
const int kConstBits = 13;
const int kPass1Bits = 2;

inline int32_t fix(double x)
{
    return (int32_t)(x * (1 << kConstBits) + 0.5);
}

inline int32_t descale(int64_t x, int n)
{
    return (int32_t)((x + ((int64_t)1 << (n - 1))) >> n);
}

void idct_islow(const int16_t* in, const uint16_t* qt, uint8_t* out, int stride)
{
    static const int32_t c0298 = fix(0.298631336), c0390 = fix(0.390180644), c0541 = fix(0.541196100),
                         c0765 = fix(0.765366865), c0899 = fix(0.899976223), c1175 = fix(1.175875602),
                         c1501 = fix(1.501321110), c1847 = fix(1.847759065), c1961 = fix(1.961570560),
                         c2053 = fix(2.053119869), c2562 = fix(2.562915447), c3072 = fix(3.072711026);
    int32_t ws[64];
    /* pass 1: columns, into ws with kPass1Bits extra bits */
    for (int col = 0; col < 8; ++col) {
        int64_t d[8];
        for (int r = 0; r < 8; ++r) {
            d[r] = (int64_t)in[r * 8 + col] * qt[r * 8 + col];
        }
        /* even part */
        int64_t z1 = (d[2] + d[6]) * c0541;
        const int64_t t2 = z1 - d[6] * c1847;
        const int64_t t3 = z1 + d[2] * c0765;
        const int64_t t0 = (d[0] + d[4]) << kConstBits;
        const int64_t t1 = (d[0] - d[4]) << kConstBits;
        const int64_t e10 = t0 + t3, e13 = t0 - t3, e11 = t1 + t2, e12 = t1 - t2;
        /* odd part */
        int64_t o0 = d[7], o1 = d[5], o2 = d[3], o3 = d[1];
        z1 = o0 + o3;
        int64_t z2 = o1 + o2, z3 = o0 + o2, z4 = o1 + o3;
        const int64_t z5 = (z3 + z4) * c1175;
        o0 *= c0298;
        o1 *= c2053;
        o2 *= c3072;
        o3 *= c1501;
        z1 *= -c0899;
        z2 *= -c2562;
        z3 *= -c1961;
        z4 *= -c0390;
        z3 += z5;
        z4 += z5;
        o0 += z1 + z3;
        o1 += z2 + z4;
        o2 += z2 + z3;
        o3 += z1 + z4;
        const int n = kConstBits - kPass1Bits;
        ws[0 * 8 + col] = descale(e10 + o3, n);
        ws[7 * 8 + col] = descale(e10 - o3, n);
        ws[1 * 8 + col] = descale(e11 + o2, n);
        ws[6 * 8 + col] = descale(e11 - o2, n);
        ws[2 * 8 + col] = descale(e12 + o1, n);
        ws[5 * 8 + col] = descale(e12 - o1, n);
        ws[3 * 8 + col] = descale(e13 + o0, n);
        ws[4 * 8 + col] = descale(e13 - o0, n);
    }
    /* pass 2: rows, removing the scale factors, level shift and clamp */
    for (int r = 0; r < 8; ++r) {
        const int32_t* w = ws + r * 8;
        int64_t z1 = ((int64_t)w[2] + w[6]) * c0541;
        const int64_t t2 = z1 - (int64_t)w[6] * c1847;
        const int64_t t3 = z1 + (int64_t)w[2] * c0765;
        const int64_t t0 = ((int64_t)w[0] + w[4]) << kConstBits;
        const int64_t t1 = ((int64_t)w[0] - w[4]) << kConstBits;
        const int64_t e10 = t0 + t3, e13 = t0 - t3, e11 = t1 + t2, e12 = t1 - t2;
        int64_t o0 = w[7], o1 = w[5], o2 = w[3], o3 = w[1];
        z1 = o0 + o3;
        int64_t z2 = o1 + o2, z3 = o0 + o2, z4 = o1 + o3;
        const int64_t z5 = (z3 + z4) * c1175;
        o0 *= c0298;
        o1 *= c2053;
        o2 *= c3072;
        o3 *= c1501;
        z1 *= -c0899;
        z2 *= -c2562;
        z3 *= -c1961;
        z4 *= -c0390;
        z3 += z5;
        z4 += z5;
        o0 += z1 + z3;
        o1 += z2 + z4;
        o2 += z2 + z3;
        o3 += z1 + z4;
        const int n = kConstBits + kPass1Bits + 3;
        const int64_t v[8] = {e10 + o3, e11 + o2, e12 + o1, e13 + o0, e13 - o0, e12 - o1, e11 - o2, e10 - o3};
        for (int c = 0; c < 8; ++c) {
            int s = descale(v[c], n) + 128;
            s = s < 0 ? 0 : s > 255 ? 255 : s;
            out[r * stride + c] = (uint8_t)s;
        }
    }
}

} // namespace

extern "C" int hpm_decode_jpeg_grey(const uint8_t* data, size_t size, uint8_t* out, int* width, int* height)
{
    if (!data || !width || !height) {
        return -1;
    }
    Decoder d;
    d.p = data;
    d.end = data + size;
    if (!d.parse()) {
        return -1;
    }
    const int W = d.width, H = d.height;
    if (!out) {
        *width = W;
        *height = H;
        return 0;
    }
    if (*width != W || *height != H) {
        return -1;
    }
    const Component& y = d.comps[0]; /* the first component is Y (or the grey channel) */
    /* the Y component covers the image at full resolution (its factors are
     * the largest in every JFIF layout); other layouts are rejected */
    if (y.h != d.hmax || y.v != d.vmax) {
        return -1;
    }
    std::vector<uint8_t> blk(64);
    for (int by = 0; by < y.bh && by * 8 < H; ++by) {
        for (int bx = 0; bx < y.bw && bx * 8 < W; ++bx) {
            idct_islow(&y.coef[((size_t)by * y.bw + bx) * 64], d.q[y.tq], blk.data(), 8);
            for (int r = 0; r < 8 && by * 8 + r < H; ++r) {
                for (int c = 0; c < 8 && bx * 8 + c < W; ++c) {
                    out[(size_t)(by * 8 + r) * W + bx * 8 + c] = blk[(size_t)r * 8 + c];
                }
            }
        }
    }
    return 0;
}

namespace {

/* one component's samples at its own (downsampled) resolution: the inverse
 * DCT of all its blocks (the padding blocks included, as libjpeg's sample
 * buffers hold them) */
std::vector<uint8_t> component_plane(const Decoder& d, const Component& c)
{
    const int pw = c.bw * 8;
    std::vector<uint8_t> plane((size_t)pw * c.bh * 8);
    for (int by = 0; by < c.bh; ++by) {
        for (int bx = 0; bx < c.bw; ++bx) {
            idct_islow(&c.coef[((size_t)by * c.bw + bx) * 64], d.q[c.tq], &plane[(size_t)by * 8 * pw + bx * 8], pw);
        }
    }
    return plane;
}

/* A chroma component at full resolution, as libjpeg's default ("fancy")
 * upsampling computes it: a triangle filter, each output sample 3/4 of the
 * nearer input sample and 1/4 of the farther one (in both directions for
 * 2x2 subsampling: 9/16, 3/16, 3/16, 1/16), with its rounding biases that
 * alternate between the two outputs of one input; the first and last
 * columns take the nearer sample alone, the rows above the first and below
 * the last real row repeat those rows; components at most two samples
 * wide are replicated instead. Factors 1x1, 2x1 and 2x2 only. */
bool upsample(const std::vector<uint8_t>& in, int pw, int cw, int ch, int hx, int vx, int W, int H,
              std::vector<uint8_t>* out)
{
    out->assign((size_t)W * H, 0);
    auto at = [&](int r, int col) { return (int)in[(size_t)r * pw + col]; };
    if (hx == 1 && vx == 1) {
        for (int y = 0; y < H; ++y) {
            for (int x = 0; x < W; ++x) {
                (*out)[(size_t)y * W + x] = (uint8_t)at(y, x);
            }
        }
        return true;
    }
    if (hx != 2 || (vx != 1 && vx != 2)) {
        return false;
    }
    if (cw <= 2) {
        /* components at most 2 samples wide are replicated (libjpeg filters
         * wider ones only) */
        for (int y = 0; y < H; ++y) {
            for (int x = 0; x < W; ++x) {
                (*out)[(size_t)y * W + x] = (uint8_t)at(y / vx, x / 2);
            }
        }
        return true;
    }
    std::vector<int> row((size_t)pw + 1), prow((size_t)2 * cw + 2);
    for (int r = 0; r < ch; ++r) {
        for (int half = 0; half < vx; ++half) {
            const int y = r * vx + half;
            if (y >= H) {
                break;
            }
            /* column sums (2x2: 3 x this row + the row above or below) */
            if (vx == 1) {
                for (int col = 0; col <= cw && col < pw; ++col) row[(size_t)col] = at(r, col);
            } else {
                const int other = half == 0 ? (r > 0 ? r - 1 : 0) : (r + 1 < ch ? r + 1 : ch - 1);
                for (int col = 0; col <= cw && col < pw; ++col) row[(size_t)col] = 3 * at(r, col) + at(other, col);
            }
            int* o = prow.data();
            {
                for (int col = 0; col < cw; ++col) {
                    const int v = row[(size_t)col];
                    int left, right;
                    if (vx == 1) {
                        left = col == 0 ? v : (v * 3 + row[(size_t)col - 1] + 1) >> 2;
                        right = col == cw - 1 ? v : (v * 3 + row[(size_t)col + 1] + 2) >> 2;
                    } else {
                        left = col == 0 ? (v * 4 + 8) >> 4 : (v * 3 + row[(size_t)col - 1] + 8) >> 4;
                        right = col == cw - 1 ? (v * 4 + 7) >> 4 : (v * 3 + row[(size_t)col + 1] + 7) >> 4;
                    }
                    o[2 * col] = left;
                    o[2 * col + 1] = right;
                }
            }
            for (int x = 0; x < W; ++x) {
                (*out)[(size_t)y * W + x] = (uint8_t)o[x];
            }
        }
    }
    return true;
}

/* JFIF YCbCr -> RGB in 16-bit fixed point with rounding, clamped to 0..255
 * (R = Y + 1.402 Cr', G = Y - 0.34414 Cb' - 0.71414 Cr', B = Y + 1.772 Cb',
 * Cb' = Cb - 128, Cr' = Cr - 128), the integer form libjpeg computes */
struct YccTables {
    int cr_r[256], cb_b[256];
    int32_t cr_g[256], cb_g[256];
    YccTables()
    {
        const int32_t half = 1 << 15;
        auto fixp = [](double x) { return (int32_t)(x * 65536.0 + 0.5); };
        for (int i = 0; i < 256; ++i) {
            const int32_t x = i - 128;
            cr_r[i] = (int)((fixp(1.40200) * x + half) >> 16);
            cb_b[i] = (int)((fixp(1.77200) * x + half) >> 16);
            cr_g[i] = -fixp(0.71414) * x;
            cb_g[i] = -fixp(0.34414) * x + half;
        }
    }
};

inline uint8_t clamp255(int v)
{
    return (uint8_t)(v < 0 ? 0 : v > 255 ? 255 : v);
}

} // namespace

extern "C" int hpm_decode_jpeg_rgb(const uint8_t* data, size_t size, uint8_t* out, int* width, int* height)
{
    if (!data || !width || !height) {
        return -1;
    }
    Decoder d;
    d.p = data;
    d.end = data + size;
    if (!d.parse()) {
        return -1;
    }
    const int W = d.width, H = d.height;
    if (!out) {
        *width = W;
        *height = H;
        return 0;
    }
    if (*width != W || *height != H) {
        return -1;
    }
    const size_t np = (size_t)W * H;
    std::vector<uint8_t> planes[3];
    for (size_t k = 0; k < d.comps.size(); ++k) {
        const Component& c = d.comps[k];
        if (d.hmax % c.h != 0 || d.vmax % c.v != 0) {
            return -1;
        }
        const int cw = (W * c.h + d.hmax - 1) / d.hmax, ch = (H * c.v + d.vmax - 1) / d.vmax;
        if (!upsample(component_plane(d, c), c.bw * 8, cw, ch, d.hmax / c.h, d.vmax / c.v, W, H, &planes[k])) {
            return -1;
        }
    }
    if (d.comps.size() == 1) {
        for (size_t i = 0; i < np; ++i) {
            out[3 * i] = out[3 * i + 1] = out[3 * i + 2] = planes[0][i];
        }
        return 0;
    }
    static const YccTables t;
    for (size_t i = 0; i < np; ++i) {
        const int y = planes[0][i], cb = planes[1][i], cr = planes[2][i];
        out[3 * i] = clamp255(y + t.cr_r[cr]);
        out[3 * i + 1] = clamp255(y + (int)((t.cb_g[cb] + t.cr_g[cr]) >> 16));
        out[3 * i + 2] = clamp255(y + t.cb_b[cb]);
    }
    return 0;
}

extern "C" int hpm_read_image_rgb(const char* path, uint8_t* data, int* width, int* height)
{
    if (!path || !width || !height) {
        return -1;
    }
    FILE* f = fopen(path, "rb");
    if (!f) {
        return -1;
    }
    std::vector<uint8_t> buf;
    uint8_t chunk[65536];
    size_t n;
    while ((n = fread(chunk, 1, sizeof(chunk), f)) > 0) {
        buf.insert(buf.end(), chunk, chunk + n);
    }
    fclose(f);
    if (buf.size() >= 2 && buf[0] == 0xFF && buf[1] == 0xD8) {
        int w = *width, h = *height;
        if (!data) {
            return hpm_decode_jpeg_rgb(buf.data(), buf.size(), nullptr, width, height);
        }
        if (hpm_decode_jpeg_rgb(buf.data(), buf.size(), nullptr, &w, &h) != 0) {
            return -1;
        }
        *width = w;
        *height = h;
        return hpm_decode_jpeg_rgb(buf.data(), buf.size(), data, &w, &h);
    }
    /* an 8-bit grey PGM: the grey value in all three channels */
    int w = 0, h = 0;
    if (hpm_read_pgm_f32(path, nullptr, &w, &h) != 0) {
        return -1;
    }
    *width = w;
    *height = h;
    if (!data) {
        return 0;
    }
    std::vector<float> g((size_t)w * h);
    if (hpm_read_pgm_f32(path, g.data(), &w, &h) != 0) {
        return -1;
    }
    for (size_t i = 0; i < g.size(); ++i) {
        data[3 * i] = data[3 * i + 1] = data[3 * i + 2] = (uint8_t)g[i];
    }
    return 0;
}

extern "C" int hpm_read_image_grey(const char* path, float* data, int* width, int* height)
{
    if (!path || !width || !height) {
        return -1;
    }
    FILE* f = fopen(path, "rb");
    if (!f) {
        return -1;
    }
    std::vector<uint8_t> buf;
    uint8_t chunk[65536];
    size_t n;
    while ((n = fread(chunk, 1, sizeof(chunk), f)) > 0) {
        buf.insert(buf.end(), chunk, chunk + n);
    }
    fclose(f);
    if (buf.size() >= 2 && buf[0] == 0xFF && buf[1] == 0xD8) {
        int w = 0, h = 0;
        if (hpm_decode_jpeg_grey(buf.data(), buf.size(), nullptr, &w, &h) != 0) {
            return -1;
        }
        if (!data) {
            *width = w;
            *height = h;
            return 0;
        }
        std::vector<uint8_t> grey((size_t)w * h);
        if (hpm_decode_jpeg_grey(buf.data(), buf.size(), grey.data(), &w, &h) != 0) {
            return -1;
        }
        for (size_t i = 0; i < grey.size(); ++i) {
            data[i] = (float)grey[i];
        }
        *width = w;
        *height = h;
        return 0;
    }
    return hpm_read_pgm_f32(path, data, width, height);
}
