"""Generate mrt_tables.h: the D3Q19 MRT transform of the reference solver
(fsi/lbm3d.py:30-54) and its exact inverse, split into the even/odd parts the
collision kernel uses.

Run once; the output is committed. The inverse is computed in exact rational
arithmetic, so every entry is the correctly rounded value of the true inverse
(the reference's table f32(inv_f64(M)) agrees except for ~1e-17 noise entries
that stand for exact zeros).
"""

from fractions import Fraction
from pathlib import Path

M = [
    [1] * 19,
    [-1, 0, 0, 0, 0, 0, 0] + [1] * 12,
    [1, -2, -2, -2, -2, -2, -2] + [1] * 12,
    [0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0],
    [0, -2, 2, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0],
    [0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1],
    [0, 0, 0, -2, 2, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1],
    [0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1],
    [0, 0, 0, 0, 0, -2, 2, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1],
    [0, 2, 2, -1, -1, -1, -1, 1, 1, 1, 1, 1, 1, 1, 1, -2, -2, -2, -2],
    [0, -2, -2, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, -2, -2, -2, -2],
    [0, 0, 0, 1, 1, -1, -1, 1, 1, 1, 1, -1, -1, -1, -1, 0, 0, 0, 0],
    [0, 0, 0, -1, -1, 1, 1, 1, 1, 1, 1, -1, -1, -1, -1, 0, 0, 0, 0],
    [0, 0, 0, 0, 0, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 0, 0, 0, 0],
    [0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, -1, -1],
    [0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0],
    [0, 0, 0, 0, 0, 0, 0, 1, -1, 1, -1, -1, 1, -1, 1, 0, 0, 0, 0],
    [0, 0, 0, 0, 0, 0, 0, -1, 1, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1],
    [0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, -1, -1, 1, -1, 1, 1, -1],
]
N = 19


def exact_inverse(a):
    n = len(a)
    aug = [[Fraction(v) for v in row] + [Fraction(int(i == j)) for j in range(n)]
           for i, row in enumerate(a)]
    for col in range(n):
        piv = next(r for r in range(col, n) if aug[r][col] != 0)
        aug[col], aug[piv] = aug[piv], aug[col]
        p = aug[col][col]
        aug[col] = [v / p for v in aug[col]]
        for r in range(n):
            if r != col and aug[r][col] != 0:
                f = aug[r][col]
                aug[r] = [vr - f * vc for vr, vc in zip(aug[r], aug[col])]
    return [row[n:] for row in aug]


def main():
    inv = exact_inverse(M)
    lr = [0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17]
    odd = [k for k in range(N) if all(M[k][q] == -M[k][lr[q]] for q in range(N))]
    for q in range(N):
        for k in range(N):
            s = -1 if k in odd else 1
            assert inv[lr[q]][k] == s * inv[q][k]
    for k in odd:
        assert inv[0][k] == 0 and M[k][0] == 0
    lines = [
        "// Generated by gen_mrt_tables.py from the D3Q19 MRT transform of the",
        "// reference solver (fsi/lbm3d.py:30-54). Do not edit.",
        "#pragma once",
        "",
        "namespace fsi_mrt {",
        "",
        "constexpr int kQ = 19;",
        "// M[k][q]: integer moment transform",
        "constexpr int M[19][19] = {",
    ]
    for row in M:
        lines.append("    {" + ", ".join(str(v) for v in row) + "},")
    lines.append("};")
    lines.append("// 1 for moments odd under e -> -e (momentum, energy flux, 3rd order)")
    lines.append("constexpr int kOdd[19] = {" + ", ".join("1" if k in odd else "0" for k in range(N)) + "};")
    lines.append("// exact inverse, as numerator/denominator pairs")
    lines.append("constexpr long long INV_NUM[19][19] = {")
    for row in inv:
        lines.append("    {" + ", ".join(str(v.numerator) for v in row) + "},")
    lines.append("};")
    lines.append("constexpr long long INV_DEN[19][19] = {")
    for row in inv:
        lines.append("    {" + ", ".join(str(v.denominator) for v in row) + "},")
    lines.append("};")
    # Guo forcing moments, fsi/lbm3d.py:276-283:
    #   G_k = sum_q M_kq w_q [(e_q - u).F + (e_q.u)(e_q.F)]
    #       = sum_a GA[k][a] F_a - GB[k] (u.F) + sum_ab GC[k][a][b] u_a F_b
    w = [Fraction(1, 3)] + [Fraction(1, 18)] * 6 + [Fraction(1, 36)] * 12
    e = [(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
         (1, 1, 0), (-1, -1, 0), (1, -1, 0), (-1, 1, 0), (1, 0, 1), (-1, 0, -1), (1, 0, -1),
         (-1, 0, 1), (0, 1, 1), (0, -1, -1), (0, 1, -1), (0, -1, 1)]
    ga = [[sum(M[k][q] * w[q] * e[q][a] for q in range(N)) for a in range(3)] for k in range(N)]
    gb = [sum(M[k][q] * w[q] for q in range(N)) for k in range(N)]
    gc = [[[sum(M[k][q] * w[q] * e[q][a] * e[q][b] for q in range(N)) for b in range(3)]
           for a in range(3)] for k in range(N)]

    def frac_table(name, flat, shape):
        lines.append(f"constexpr long long {name}_NUM{shape} = {{" + ", ".join(str(v.numerator) for v in flat) + "};")
        lines.append(f"constexpr long long {name}_DEN{shape} = {{" + ", ".join(str(v.denominator) for v in flat) + "};")

    # M^-1 = M^T (M M^T)^-1; the Gram matrix M M^T only couples moments of the
    # same parity and order, so its inverse is block diagonal and cheap to apply
    gram = [[sum(M[a][q] * M[b][q] for q in range(N)) for b in range(N)] for a in range(N)]
    ginv = exact_inverse(gram)
    for a in range(N):
        for b in range(N):
            assert sum(M[q][a] * ginv[q][b] for q in range(N)) == inv[a][b]
    lines.append("// inverse Gram matrix (M M^T)^-1, block diagonal: M^-1 = M^T GINV")
    lines.append("constexpr long long GINV_NUM[19][19] = {")
    for row in ginv:
        lines.append("    {" + ", ".join(str(v.numerator) for v in row) + "},")
    lines.append("};")
    lines.append("constexpr long long GINV_DEN[19][19] = {")
    for row in ginv:
        lines.append("    {" + ", ".join(str(v.denominator) for v in row) + "},")
    lines.append("};")
    lines.append("// Guo moment tables, flattened: GA[k*3+a], GB[k], GC[(k*3+a)*3+b]")
    frac_table("GA", [ga[k][a] for k in range(N) for a in range(3)], "[57]")
    frac_table("GB", gb, "[19]")
    frac_table("GC", [gc[k][a][b] for k in range(N) for a in range(3) for b in range(3)], "[171]")
    lines.append("")
    lines.append("}  // namespace fsi_mrt")
    out = Path(__file__).with_name("mrt_tables.h")
    out.write_text("\n".join(lines) + "\n")
    print("wrote", out)


if __name__ == "__main__":
    main()
