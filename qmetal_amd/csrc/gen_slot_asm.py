"""Writes slot_asm.inc: one inline-assembly statement per 2x2 slot op of a
merged pass, for tiles of 2^NL amplitudes (NL = 4, 5).

    python qmetal_amd/csrc/gen_slot_asm.py      (build.py runs it)

Why assembly: a pass applies a runtime list of ops to the tile held in
registers. Written in C++, every op kind or target bit is a separate code path
that redefines the tile, and the compiler then keeps a second copy of the
whole tile across the op loop (double the VGPRs, half the occupancy). One
statement whose outputs are tied to its inputs defines every amplitude exactly
once per op, whatever section runs inside it.

Section `sel` = 8 * kind + lbit:
  kind 0  real matrix (m00, m01, m10, m11 = m[0], m[2], m[4], m[6]), all pairs
          (8 instructions per pair)
  kind 1  complex matrix, all pairs (16 per pair)
  kind 2  complex matrix, only the pairs whose bit is set in `pm` (pair p is
          the p-th index l with bit lbit clear, ascending)
  kind 5  X: swap the two amplitudes of the pairs in `pm` (2 per pair)
  kind 6  diagonal matrix diag(m00, m11) on the pairs in `pm` (8 per pair)
  kind 7  real diagonal, imaginary off-diagonal (RX, and products of RX
          with real phases: m00, m11 = m[0], m[6]; m01, m10 = i m[3], i m[5]),
          all pairs (8 per pair)
Any other `sel` leaves the tile unchanged. The matrix is per lane (VGPRs).
(Kinds 3 and 4, complex matrices on the matrix core, were measured and
dropped: slower than the VALU in these passes.)
"""
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def bfly_real(a, b):
    """(a, b) <- (m00 a + m01 b, m10 a + m11 b), real m: each output's first
    product goes to a temporary, the second is a fused multiply-add written
    in place as the input's last use (no moves)."""
    ra, ia, rb, ib = f"%[r{a}]", f"%[i{a}]", f"%[r{b}]", f"%[i{b}]"
    out = []
    for xa, xb in ((ra, rb), (ia, ib)):
        out += [
            f"v_mul_f32 %[t0], %[m2r], {xa}",
            f"v_mul_f32 %[t1], %[m1r], {xb}",
            f"v_fma_f32 {xb}, %[m3r], {xb}, %[t0]",
            f"v_fma_f32 {xa}, %[m0r], {xa}, %[t1]",
        ]
    return out


def bfly_cplx(a, b):
    """Complex m: three terms of each output into four temporaries, then the
    fourth term of each, fused and written in place (no moves)."""
    ra, ia, rb, ib = f"%[r{a}]", f"%[i{a}]", f"%[r{b}]", f"%[i{b}]"
    return [
        # a'.x = m00r ax - m00i ay + m01r bx - m01i by
        f"v_mul_f32 %[t0], %[m0i], {ia}",
        f"v_fma_f32 %[t0], %[m1r], {rb}, -%[t0]",
        f"v_fma_f32 %[t0], -%[m1i], {ib}, %[t0]",
        # a'.y = m00r ay + m00i ax + m01r by + m01i bx
        f"v_mul_f32 %[t1], %[m0i], {ra}",
        f"v_fmac_f32 %[t1], %[m1r], {ib}",
        f"v_fmac_f32 %[t1], %[m1i], {rb}",
        # b'.x = m10r ax - m10i ay + m11r bx - m11i by
        f"v_mul_f32 %[t2], %[m2r], {ra}",
        f"v_fma_f32 %[t2], -%[m2i], {ia}, %[t2]",
        f"v_fma_f32 %[t2], -%[m3i], {ib}, %[t2]",
        # b'.y = m10r ay + m10i ax + m11r by + m11i bx
        f"v_mul_f32 %[t3], %[m2r], {ia}",
        f"v_fmac_f32 %[t3], %[m2i], {ra}",
        f"v_fmac_f32 %[t3], %[m3i], {rb}",
        # last terms, in place
        f"v_fma_f32 {ra}, %[m0r], {ra}, %[t0]",
        f"v_fma_f32 {ia}, %[m0r], {ia}, %[t1]",
        f"v_fma_f32 {rb}, %[m3r], {rb}, %[t2]",
        f"v_fma_f32 {ib}, %[m3r], {ib}, %[t3]",
    ]


def bfly_rx(a, b):
    """(a, b) <- (m00 a + i m01i b, i m10i a + m11 b) with m00, m11 real: the
    four cross products into temporaries, then fused in-place updates."""
    ra, ia, rb, ib = f"%[r{a}]", f"%[i{a}]", f"%[r{b}]", f"%[i{b}]"
    return [
        f"v_mul_f32 %[t0], %[m1i], {ib}",
        f"v_mul_f32 %[t1], %[m1i], {rb}",
        f"v_mul_f32 %[t2], %[m2i], {ia}",
        f"v_mul_f32 %[t3], %[m2i], {ra}",
        f"v_fma_f32 {ra}, %[m0r], {ra}, -%[t0]",
        f"v_fma_f32 {ia}, %[m0r], {ia}, %[t1]",
        f"v_fma_f32 {rb}, %[m3r], {rb}, -%[t2]",
        f"v_fma_f32 {ib}, %[m3r], {ib}, %[t3]",
    ]


def pairs(nl, h):
    return [(l, l | (1 << h)) for l in range(1 << nl) if not (l >> h) & 1]


def swap_section(nl, h):
    out = []
    for p, (a, b) in enumerate(pairs(nl, h)):
        out += [f"s_bitcmp1_b32 %[pm], {p}", f"s_cbranch_scc0 .Lqx%=_{h}_{p}",
                f"v_swap_b32 %[r{a}], %[r{b}]", f"v_swap_b32 %[i{a}], %[i{b}]",
                f".Lqx%=_{h}_{p}:"]
    return out


def diag_section(nl, h):
    """a <- m00 a, b <- m11 b (complex), on the pairs in pm: two products
    into temporaries, then the fused in-place updates (no moves)."""
    out = []
    for p, (a, b) in enumerate(pairs(nl, h)):
        out += [f"s_bitcmp1_b32 %[pm], {p}", f"s_cbranch_scc0 .Lqd%=_{h}_{p}"]
        for x, mr, mi in ((a, "%[m0r]", "%[m0i]"), (b, "%[m3r]", "%[m3i]")):
            rx, ix = f"%[r{x}]", f"%[i{x}]"
            out += [f"v_mul_f32 %[t0], {mi}, {ix}",
                    f"v_mul_f32 %[t1], {mi}, {rx}",
                    f"v_fma_f32 {rx}, {mr}, {rx}, -%[t0]",
                    f"v_fma_f32 {ix}, {mr}, {ix}, %[t1]"]
        out.append(f".Lqd%=_{h}_{p}:")
    return out


KIND_ORDER = (1, 6, 7, 5, 0, 2)  # dispatch order: the common kinds first


def statement(nl):
    lines = ["s_lshr_b32 %[kk], %[sel], 3", "s_and_b32 %[hh], %[sel], 7"]
    for kind in KIND_ORDER:
        lines += [f"s_cmp_eq_u32 %[kk], {kind}", f"s_cbranch_scc1 .Lqk%=_{kind}"]
    lines.append("s_branch .Lqs%=_end")
    for kind in KIND_ORDER:
        lines.append(f".Lqk%=_{kind}:")
        for h in range(nl):
            lines += [f"s_cmp_eq_u32 %[hh], {h}", f"s_cbranch_scc1 .Lqs%=_{kind}_{h}"]
        lines.append("s_branch .Lqs%=_end")
        for h in range(nl):
            lines.append(f".Lqs%=_{kind}_{h}:")
            if kind == 5:
                lines += swap_section(nl, h)
            elif kind == 6:
                lines += diag_section(nl, h)
            else:
                for p, (a, b) in enumerate(pairs(nl, h)):
                    if kind == 0:
                        lines += bfly_real(a, b)
                    elif kind == 1:
                        lines += bfly_cplx(a, b)
                    elif kind == 7:
                        lines += bfly_rx(a, b)
                    else:
                        lines += [f"s_bitcmp1_b32 %[pm], {p}", f"s_cbranch_scc0 .Lqs%=_{kind}_{h}_{p}"]
                        lines += bfly_cplx(a, b)
                        lines.append(f".Lqs%=_{kind}_{h}_{p}:")
            lines.append("s_branch .Lqs%=_end")
    lines.append(".Lqs%=_end:")
    body = "\n".join(f'      "{ln}\\n"' for ln in lines)
    n = 1 << nl
    outs = ",\n".join(
        f'        [r{k}] "+v"(x[{k}].x), [i{k}] "+v"(x[{k}].y)' for k in range(n))
    return f"""template <>
__device__ __forceinline__ void slot_asm<{nl}>(float2 (&x)[{n}], uint32_t sel, uint32_t pm,
                                             const float (&m)[8]) {{
  float t0, t1, t2, t3;
  uint32_t kk, hh;
  asm(
{body}
      : {outs.strip()},
        [t0] "=&v"(t0), [t1] "=&v"(t1), [t2] "=&v"(t2), [t3] "=&v"(t3),
        [kk] "=&s"(kk), [hh] "=&s"(hh)
      : [m0r] "v"(m[0]), [m0i] "v"(m[1]), [m1r] "v"(m[2]), [m1i] "v"(m[3]),
        [m2r] "v"(m[4]), [m2i] "v"(m[5]), [m3r] "v"(m[6]), [m3i] "v"(m[7]),
        [sel] "s"(sel), [pm] "s"(pm)
      : "scc");
}}
"""


def main():
    out = [
        "// slot_asm.inc -- generated by gen_slot_asm.py; do not edit.",
        "// One inline-assembly statement per 2x2 slot op of a merged pass (see the",
        "// generator for the section layout and why the op is written this way).",
        "",
        "template <int NL>",
        "__device__ __forceinline__ void slot_asm(float2 (&x)[1 << NL], uint32_t sel, uint32_t pm,",
        "                                         const float (&m)[8]);",
        "",
    ]
    for nl in (4, 5):
        out.append(statement(nl))
    text = "\n".join(out)
    path = os.path.join(HERE, "slot_asm.inc")
    old = open(path).read() if os.path.exists(path) else None
    if old != text:
        with open(path, "w") as f:
            f.write(text)
    return path


if __name__ == "__main__":
    print(main())
