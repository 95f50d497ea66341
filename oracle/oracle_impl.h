/*
 * ORACLE — type-generic half of waltz_oracle.c (test infrastructure only).
 * Included twice with T = float / double. Requires: T, SFX (f32/f64), FMA,
 * FABS, LLRINT, WIDTH. See waltz_oracle.h for the pinning of each function.
 */
#define OCAT2(a, b) a##_##b
#define OCAT(a, b) OCAT2(a, b)
#define FN(name) OCAT(name, SFX)

/* Lifting constants, cpu/src/transform.cpp:41-48 (rounded to T; ik divided in T). */
#define LA ((T)-1.586134342059924)
#define LB ((T)-0.052980118572961)
#define LG ((T)0.882911075530934)
#define LD ((T)0.443506852043971)
#define LK ((T)1.1496043988602418)

/* One line, tiled halo lifting: cpu/src/transform.cpp:113-179.
 * in: the line's samples (forward) or split coefficients (inverse), length n.
 * out: written at split positions (forward) or sample positions (inverse). */
static void FN(lift_line)(const T* in, T* out, uint32_t n, int tile, int inverse, int folded) {
    const int pairs = tile / 2 + 4;
    const T ik = (T)1 / LK;
    T e[40], o[40];
    for (uint64_t start = 0; start < n; start += (unsigned)tile) {
        for (int j = 0; j < pairs; ++j) {
            const int64_t p = (int64_t)start - 4 + 2 * j;
            e[j] = in[inverse ? split_pos(p, n) : reflect_pos(p, n)];
            o[j] = in[inverse ? split_pos(p + 1, n) : reflect_pos(p + 1, n)];
        }
        if (inverse) {
            const T first_odd = o[0], last_odd = o[pairs - 1];
            for (int j = 0; j < pairs; ++j) { e[j] *= ik; o[j] *= -LK; }
            for (int j = 1; j < pairs; ++j) {
                T s;
                if (folded && j == 1) s = FMA(first_odd, -LK, o[1]);
                else if (folded && j == pairs - 1) s = FMA(last_odd, -LK, o[pairs - 2]);
                else s = o[j - 1] + o[j];
                e[j] = FMA(-LD, s, e[j]);
            }
            for (int j = 1; j < pairs - 1; ++j) o[j] = FMA(-LG, e[j] + e[j + 1], o[j]);
            for (int j = 2; j < pairs - 1; ++j) e[j] = FMA(-LB, o[j - 1] + o[j], e[j]);
            for (int j = 2; j < pairs - 2; ++j) o[j] = FMA(-LA, e[j] + e[j + 1], o[j]);
        } else {
            for (int j = 0; j < pairs - 1; ++j) o[j] = FMA(LA, e[j] + e[j + 1], o[j]);
            for (int j = 1; j < pairs - 1; ++j) e[j] = FMA(LB, o[j - 1] + o[j], e[j]);
            for (int j = 1; j < pairs - 2; ++j) o[j] = FMA(LG, e[j] + e[j + 1], o[j]);
            for (int j = 2; j < pairs - 2; ++j) e[j] = FMA(LD, o[j - 1] + o[j], e[j]);
            for (int j = 2; j < pairs - 2; ++j) { e[j] *= LK; o[j] *= -ik; }
        }
        for (int j = 2; j < pairs - 2; ++j) {
            const uint64_t p = start + 2 * (uint64_t)(j - 2);
            if (p < n) out[inverse ? p : p / 2] = e[j];
            if (p + 1 < n) out[inverse ? p + 1 : half_len(n) + p / 2] = o[j];
        }
    }
}

/* One axis over the active box [0,a.nx)x[0,a.ny)x[0,a.nz) of a full volume:
 * cpu/src/transform.cpp:88-108,149-179. Lines are independent; each is staged
 * whole before it is written back. */
static void FN(lift_axis)(T* data, T* line_in, T* line_out, oshape full, oshape a, int axis, int tile,
                          int inverse) {
    const uint32_t n = axis == 0 ? a.nx : axis == 1 ? a.ny : a.nz;
    const size_t stride = axis == 0 ? 1 : axis == 1 ? full.nx : (size_t)full.nx * full.ny;
    const uint32_t u_len = axis == 0 ? a.ny : a.nx;           /* first free coordinate */
    const uint32_t v_len = axis == 2 ? a.ny : a.nz;           /* second free coordinate */
    for (uint32_t v = 0; v < v_len; ++v)
        for (uint32_t u = 0; u < u_len; ++u) {
            size_t base;
            if (axis == 0) base = ((size_t)v * full.ny + u) * full.nx;
            else if (axis == 1) base = (size_t)v * full.nx * full.ny + u;
            else base = (size_t)v * full.nx + u;
            for (uint32_t p = 0; p < n; ++p) line_in[p] = data[base + p * stride];
            FN(lift_line)(line_in, line_out, n, tile, inverse, axis != 0);
            for (uint32_t p = 0; p < n; ++p) data[base + p * stride] = line_out[p];
        }
}

/* Level schedules, cpu/src/transform.cpp:409-456. Tiles: forward Y 8, X 56,
 * Z 8 (f32) / 4 (f64) dyadic, Z 16/8 plane; inverse Z 16/8, X 56, Y 8. */
static int FN(transform_core)(T* data, oshape full, unsigned lxy, unsigned lz, int dyadic, int inverse) {
    if (lxy > 5 || lz > 5 || (dyadic && lxy != lz)) return fail("invalid transform levels");
    const uint32_t longest = full.nx > full.ny ? (full.nx > full.nz ? full.nx : full.nz)
                                               : (full.ny > full.nz ? full.ny : full.nz);
    T* line = (T*)malloc(sizeof(T) * 2 * (size_t)longest);
    if (!line) return fail("out of memory");
    T* line_out = line + longest;
    const int ztile = WIDTH == 4 ? 16 : 8;
    if (dyadic) {
        oshape sizes[5], act = full;
        for (unsigned l = 0; l < lxy; ++l) { sizes[l] = act; act = shrink(act, 1, 1); }
        for (unsigned step = 0; step < lxy; ++step) {
            const unsigned level = inverse ? lxy - 1 - step : step;
            act = sizes[level];
            if (inverse) {
                FN(lift_axis)(data, line, line_out, full, act, 2, ztile, 1);
                FN(lift_axis)(data, line, line_out, full, act, 0, 56, 1);
                FN(lift_axis)(data, line, line_out, full, act, 1, 8, 1);
            } else {
                FN(lift_axis)(data, line, line_out, full, act, 1, 8, 0);
                FN(lift_axis)(data, line, line_out, full, act, 0, 56, 0);
                FN(lift_axis)(data, line, line_out, full, act, 2, WIDTH == 4 ? 8 : 4, 0);
            }
        }
    } else {
        oshape xy[5], zs[5], act = full;
        for (unsigned l = 0; l < lxy; ++l) { xy[l] = act; act = shrink(act, 1, 0); }
        act = full;
        for (unsigned l = 0; l < lz; ++l) { zs[l] = act; act = shrink(act, 0, 1); }
        if (!inverse) {
            for (unsigned l = 0; l < lz; ++l) FN(lift_axis)(data, line, line_out, full, zs[l], 2, ztile, 0);
            for (unsigned l = 0; l < lxy; ++l) {
                FN(lift_axis)(data, line, line_out, full, xy[l], 1, 8, 0);
                FN(lift_axis)(data, line, line_out, full, xy[l], 0, 56, 0);
            }
        } else {
            for (unsigned s = 0; s < lxy; ++s) {
                const unsigned l = lxy - 1 - s;
                FN(lift_axis)(data, line, line_out, full, xy[l], 0, 56, 1);
                FN(lift_axis)(data, line, line_out, full, xy[l], 1, 8, 1);
            }
            for (unsigned s = 0; s < lz; ++s) FN(lift_axis)(data, line, line_out, full, zs[lz - 1 - s], 2, ztile, 1);
        }
    }
    free(line);
    return 0;
}

int FN(wo_transform)(T* data, uint32_t nx, uint32_t ny, uint32_t nz, unsigned lxy, unsigned lz, int dyadic,
                     int inverse) {
    if (!data || !nx || !ny || !nz) return fail("invalid transform arguments");
    fpguard g;
    guard_enter(&g);
    const oshape s = {nx, ny, nz};
    const int rc = FN(transform_core)(data, s, lxy, lz, dyadic, inverse);
    guard_leave(&g);
    return rc;
}

/* Quantization in Morton brick order: cpu/src/codec.cpp:198-234 (scalar rule,
 * also cpu/src/quantize_avx2.cpp:45-62). */
static int FN(quantize_core)(T* coeff, oshape s, uint32_t bz, double quantum, uint16_t* mag, uint32_t* sign,
                             uint32_t* qidx, uint32_t* qval, uint32_t* nq_out) {
    const T q = (T)quantum, inv = (T)(1.0 / quantum);
    const uint32_t n = s.nx * s.ny * s.nz;
    const size_t plane = (size_t)s.nx * s.ny;
    uint16_t* rank = brick_ranks(s, bz);
    if (!rank) return fail("out of memory");
    memset(sign, 0, ((size_t)n + 31) / 32 * 4);
    uint32_t nq = 0;
    size_t base = 0;
    int bad = 0;
    for (uint32_t z0 = 0; z0 < s.nz && !bad; z0 += bz)
        for (uint32_t y0 = 0; y0 < s.ny && !bad; y0 += 16)
            for (uint32_t x0 = 0; x0 < s.nx && !bad; x0 += 16) {
                const uint32_t bx = min_u(16, s.nx - x0), by = min_u(16, s.ny - y0), bzz = min_u(bz, s.nz - z0);
                const unsigned shp = (bx != 16) | ((by != 16) << 1) | ((bzz != bz) << 2);
                const uint16_t* r = rank + shp * 8192;
                for (uint32_t z = 0; z < bzz && !bad; ++z)
                    for (uint32_t y = 0; y < by && !bad; ++y)
                        for (uint32_t x = 0; x < bx; ++x) {
                            const size_t at = (size_t)(z0 + z) * plane + (size_t)(y0 + y) * s.nx + x0 + x;
                            const size_t packed = base + r[(z * by + y) * bx + x];
                            const T scaled = coeff[at] * inv;
                            if (!isfinite(scaled) || FABS(scaled) >= (T)9223372036854775808.0) { bad = 1; break; }
                            const long long v = LLRINT(scaled);
                            const int negative = v < 0;
                            const uint64_t m64 = (uint64_t)(negative ? -v : v);
                            const int32_t m = (int32_t)(m64 > INT32_MAX ? INT32_MAX : m64);
                            coeff[at] = m ? (T)(negative ? -m : m) * q : (T)0;
                            mag[packed] = (uint16_t)(m > 65535 ? 65535 : m);
                            if (negative && m) sign[packed / 32] |= 1u << (packed % 32);
                            if (m > 65535) { qidx[nq] = (uint32_t)at; qval[nq++] = (uint32_t)m; }
                        }
                base += (size_t)bx * by * bzz;
            }
    free(rank);
    if (bad) return fail("coefficient exceeds portable integer conversion domain");
    sort_index_pairs(qidx, qval, nq);
    *nq_out = nq;
    return 0;
}

int FN(wo_quantize)(T* coeff, uint32_t nx, uint32_t ny, uint32_t nz, uint32_t bz, double quantum, uint16_t* mag,
                    uint32_t* sign, uint32_t* qidx, uint32_t* qval, uint32_t* nq) {
    fpguard g;
    guard_enter(&g);
    const oshape s = {nx, ny, nz};
    const int rc = FN(quantize_core)(coeff, s, bz, quantum, mag, sign, qidx, qval, nq);
    guard_leave(&g);
    return rc;
}

/* Packed -> row-major dequantized, cpu/src/codec.cpp:139-150. */
static int FN(dequantize_core)(const uint16_t* mag, const uint32_t* sign, oshape s, uint32_t bz, double quantum,
                               T* out) {
    const T q = (T)quantum;
    const size_t plane = (size_t)s.nx * s.ny;
    uint16_t* rank = brick_ranks(s, bz);
    if (!rank) return fail("out of memory");
    size_t base = 0;
    for (uint32_t z0 = 0; z0 < s.nz; z0 += bz)
        for (uint32_t y0 = 0; y0 < s.ny; y0 += 16)
            for (uint32_t x0 = 0; x0 < s.nx; x0 += 16) {
                const uint32_t bx = min_u(16, s.nx - x0), by = min_u(16, s.ny - y0), bzz = min_u(bz, s.nz - z0);
                const unsigned shp = (bx != 16) | ((by != 16) << 1) | ((bzz != bz) << 2);
                const uint16_t* r = rank + shp * 8192;
                for (uint32_t z = 0; z < bzz; ++z)
                    for (uint32_t y = 0; y < by; ++y)
                        for (uint32_t x = 0; x < bx; ++x) {
                            const size_t packed = base + r[(z * by + y) * bx + x];
                            const T v = (T)mag[packed] * q;
                            out[(size_t)(z0 + z) * plane + (size_t)(y0 + y) * s.nx + x0 + x] =
                                (sign[packed / 32] >> (packed % 32) & 1u) ? -v : v;
                        }
                base += (size_t)bx * by * bzz;
            }
    free(rank);
    return 0;
}

void FN(wo_dequantize)(const uint16_t* mag, const uint32_t* sign, uint32_t nx, uint32_t ny, uint32_t nz, uint32_t bz,
                       double quantum, T* out) {
    fpguard g;
    guard_enter(&g);
    const oshape s = {nx, ny, nz};
    FN(dequantize_core)(mag, sign, s, bz, quantum, out);
    guard_leave(&g);
}

/* Native repairs, cpu/src/repair_avx2.cpp:58-74. */
static int FN(repairs_core)(const T* original, const T* rec, uint32_t n, double bound, uint32_t* idx, T* err,
                            uint32_t* count) {
    const T native_bound = (T)bound;
    uint32_t c = 0;
    for (uint32_t i = 0; i < n; ++i) {
        T value = rec[i];
        const T e = value - original[i];
        if (!isfinite(value) || !isfinite(e)) { *count = c; return fail("nonfinite reconstruction"); }
        if (FABS(e) > native_bound) {
            idx[c] = i;
            err[c++] = e;
            value = value - e;
        }
        if (!isfinite(value) || !wo_within_bound((double)value, (double)original[i], bound)) {
            *count = c;
            return fail("native repair cannot meet strict bound");
        }
    }
    *count = c;
    return 0;
}

int FN(wo_repairs)(const T* original, const T* reconstructed, uint32_t n, double bound, uint32_t* idx, T* err,
                   uint32_t* count) {
    fpguard g;
    guard_enter(&g);
    const int rc = FN(repairs_core)(original, reconstructed, n, bound, idx, err, count);
    guard_leave(&g);
    return rc;
}

static void FN(put_real)(uint8_t* p, T v) {
    if (WIDTH == 4) { uint32_t u; memcpy(&u, &v, 4); put32(p, u); }
    else { uint64_t u; memcpy(&u, &v, 8); put64(p, u); }
}
static T FN(get_real)(const uint8_t* p) {
    T v;
    if (WIDTH == 4) { uint32_t u = get32(p); memcpy(&v, &u, 4); }
    else { uint64_t u = get64(p); memcpy(&v, &u, 8); }
    return v;
}

/* Encoder, cpu/src/codec.cpp:177-258. */
static int FN(encode)(const T* input, oshape s, double abs_bound, uint8_t* out, size_t cap, size_t* written) {
    const uint32_t n = s.nx * s.ny * s.nz;
    const unsigned lx = wo_dwt_levels(min_u(s.nx, s.ny)), lz = wo_dwt_levels(s.nz);
    const int dyadic = lx == lz;
    const uint32_t bz = wo_zband(s.nx, s.ny, s.nz, lz, dyadic);
    const double quantum = 1.5 * abs_bound;
    const T q = (T)quantum, inv = (T)(1.0 / quantum);
    if (!(q > (T)0) || !isfinite(q) || !isfinite(inv)) return fail("quantum not representable in native type");
    int rc = -1;
    T* data = (T*)malloc(sizeof(T) * (size_t)n);
    T* err = (T*)malloc(sizeof(T) * (size_t)n);
    uint16_t* mag = (uint16_t*)malloc(2 * (size_t)n);
    uint32_t* sign = (uint32_t*)malloc(((size_t)n + 31) / 32 * 4);
    uint32_t* qidx = (uint32_t*)malloc(4 * (size_t)n);
    uint32_t* qval = (uint32_t*)malloc(4 * (size_t)n);
    uint32_t* pidx = (uint32_t*)malloc(4 * (size_t)n);
    if (!data || !err || !mag || !sign || !qidx || !qval || !pidx) { fail("out of memory"); goto done; }
    for (uint32_t i = 0; i < n; ++i) {
        if (!isfinite(input[i])) { fail("nonfinite input is unsupported"); goto done; }
        data[i] = input[i];
    }
    if (FN(transform_core)(data, s, lx, lz, dyadic, 0)) goto done;
    uint32_t nq = 0, np = 0;
    if (FN(quantize_core)(data, s, bz, quantum, mag, sign, qidx, qval, &nq)) goto done;
    if (FN(transform_core)(data, s, lx, lz, dyadic, 1)) goto done;
    if (FN(repairs_core)(input, data, n, abs_bound, pidx, err, &np)) goto done;
    if ((n > (1u << 20) && nq >= (1u << 20)) || (n > (1u << 22) && np >= (1u << 22))) {
        fail("portable GPU side-channel capacity reached");
        goto done;
    }
    if (cap < 32) { fail("archive capacity too small"); goto done; }
    const int64_t wb = wo_wzp_encode(mag, sign, n, out + 32, cap - 32);
    if (wb < 0) goto done;
    const blob_layout L = layout((uint64_t)wb, np, np * WIDTH, nq);
    if (L.end > cap) { fail("archive capacity too small"); goto done; }
    out[0] = (uint8_t)(dyadic | (WIDTH == 8 ? 4 : 0) | (3 << 3));
    out[1] = (uint8_t)lx;
    out[2] = (uint8_t)lz;
    out[3] = (uint8_t)bz;
    put32(out + 4, np * WIDTH);
    { uint64_t qb; memcpy(&qb, &quantum, 8); put64(out + 8, qb); }
    put64(out + 16, (uint64_t)wb);
    put32(out + 24, np);
    put32(out + 28, nq);
    memset(out + 32 + wb, 0, L.end - 32 - (size_t)wb);
    for (uint32_t i = 0; i < np; ++i) {
        put32(out + L.pi + 4 * (size_t)i, pidx[i]);
        FN(put_real)(out + L.pv + WIDTH * (size_t)i, err[i]);
    }
    for (uint32_t i = 0; i < nq; ++i) {
        put32(out + L.qi + 4 * (size_t)i, qidx[i]);
        put32(out + L.qv + 4 * (size_t)i, qval[i]);
    }
    *written = (size_t)L.end;
    rc = 0;
done:
    free(data); free(err); free(mag); free(sign); free(qidx); free(qval); free(pidx);
    return rc;
}

int FN(wo_compress)(const T* in, uint32_t nx, uint32_t ny, uint32_t nz, double abs_bound, uint8_t* out,
                    size_t capacity, size_t* written) {
    if (written) *written = 0;
    if (!in || !out || !written) return fail("null argument");
    if (!(abs_bound > 0) || !isfinite(abs_bound)) return fail("positive finite absolute error required");
    if (!nx || !ny || !nz) return fail("nonzero dimensions required");
    const uint64_t n = (uint64_t)nx * ny * nz;
    if (n > UINT32_MAX / WIDTH) return fail("volume exceeds WLT2 section limits");
    fpguard g;
    guard_enter(&g);
    const oshape s = {nx, ny, nz};
    const int rc = FN(encode)(in, s, abs_bound, out, capacity, written);
    guard_leave(&g);
    return rc;
}

/* Decoder, cpu/src/codec.cpp:260-336 (native repairs only; the compact WPW1
 * repair form is a CUDA-only profile that the canonical encoder never emits). */
static int FN(decode)(const uint8_t* in, size_t size, oshape s, T* out) {
    const uint32_t n = s.nx * s.ny * s.nz;
    if (size < 32 || (in[0] & 0xe0) || !(in[0] & 0x18) || ((in[0] & 4) ? 8 : 4) != WIDTH)
        return fail("invalid Waltz archive header");
    if (((in[0] >> 3) & 3) != 3) return fail("unsupported Waltz synthesis mode");
    if (in[0] & 2) return fail("compact repairs are not part of the canonical profile");
    const unsigned lx = in[1], lz = in[2], bz = in[3];
    const int dyadic = in[0] & 1;
    if (lx > 5 || lz > 5 || bz < 1 || bz > 32 || (dyadic && lx != lz)) return fail("invalid levels or z band");
    if (lx != wo_dwt_levels(min_u(s.nx, s.ny)) || lz != wo_dwt_levels(s.nz)) return fail("dimensions/levels mismatch");
    double quantum;
    { uint64_t qb = get64(in + 8); memcpy(&quantum, &qb, 8); }
    const T q = (T)quantum;
    if (!(quantum > 0) || !isfinite(quantum) || !(q > (T)0) || !isfinite(q)) return fail("invalid quantum");
    const uint64_t w64 = get64(in + 16);
    const uint32_t np = get32(in + 24), nq = get32(in + 28), pb = get32(in + 4);
    if (np > n || nq > n) return fail("invalid section counts");
    const blob_layout L = layout(w64, np, pb, nq);
    if (L.end != size) return fail("truncated archive or trailing data");
    if ((uint64_t)pb != (uint64_t)np * WIDTH) return fail("invalid native correction section");
    int rc = -1;
    T* data = (T*)malloc(sizeof(T) * (size_t)n);
    uint16_t* mag = (uint16_t*)malloc(2 * (size_t)n);
    uint32_t* sign = (uint32_t*)malloc(((size_t)n + 31) / 32 * 4);
    if (!data || !mag || !sign) { fail("out of memory"); goto done; }
    if (wo_wzp_decode(in + 32, (size_t)w64, n, mag, sign)) goto done;
    if (FN(dequantize_core)(mag, sign, s, bz, quantum, data)) goto done;
    if (check_indices(in + L.qi, nq, n)) goto done;
    for (uint32_t i = 0; i < nq; ++i) {
        const uint32_t idx = get32(in + L.qi + 4 * (size_t)i), v = get32(in + L.qv + 4 * (size_t)i);
        if (v <= 65535 || v > INT32_MAX) { fail("invalid quantization outlier"); goto done; }
        const T x = (T)v * q;
        data[idx] = data[idx] < (T)0 ? -x : x;
    }
    if (check_indices(in + L.pi, np, n)) goto done;
    for (uint32_t i = 0; i < np; ++i)
        if (!isfinite(FN(get_real)(in + L.pv + WIDTH * (size_t)i))) { fail("nonfinite correction"); goto done; }
    if (FN(transform_core)(data, s, lx, lz, dyadic, 1)) goto done;
    for (uint32_t i = 0; i < np; ++i) {
        const uint32_t idx = get32(in + L.pi + 4 * (size_t)i);
        data[idx] = data[idx] - FN(get_real)(in + L.pv + WIDTH * (size_t)i);
    }
    for (uint32_t i = 0; i < n; ++i)
        if (!isfinite(data[i])) { fail("nonfinite reconstruction"); goto done; }
    memcpy(out, data, sizeof(T) * (size_t)n);
    rc = 0;
done:
    free(data); free(mag); free(sign);
    return rc;
}

int FN(wo_decompress)(const uint8_t* in, size_t bytes, uint32_t nx, uint32_t ny, uint32_t nz, T* out) {
    if (!in || !out) return fail("null argument");
    if (!nx || !ny || !nz) return fail("nonzero dimensions required");
    const uint64_t n = (uint64_t)nx * ny * nz;
    if (n > UINT32_MAX / WIDTH) return fail("volume exceeds WLT2 section limits");
    fpguard g;
    guard_enter(&g);
    const oshape s = {nx, ny, nz};
    const int rc = FN(decode)(in, bytes, s, out);
    guard_leave(&g);
    return rc;
}

#undef LA
#undef LB
#undef LG
#undef LD
#undef LK
#undef FN
#undef OCAT
#undef OCAT2
