"""Writes tests/golden/hashgrid_levels.json: known answers for the hash-grid level
table of models/encoders.py:76-103 (b, resolutions, 8-aligned dense sizes,
first_hash_level, total rows).

Computed here with 50-digit decimal arithmetic (independent of the float64 host
code under test), and checked against the values quoted for the reference's
configurations: L=16, T=2^19, N in [16, 2048] -> b = 1.381913, first_hash_level 5,
dense sizes 4096/12168/29792/79512/205384, 6,098,120 rows; T=2^22 -> first_hash_level
8, 39,474,024 rows; L=5, T=2^17, N in [16, 512] -> b = 2.378414, first_hash_level 2,
456,632 rows.
"""
import json
from decimal import Decimal, getcontext, ROUND_CEILING
from pathlib import Path

getcontext().prec = 50


def level_table(L, T, nmin, nmax):
    b = (Decimal(nmax).ln() - Decimal(nmin).ln()) / (L - 1)
    b = b.exp()
    res, sizes, first = [], [], 0
    for i in range(L):
        s = Decimal(nmin) * b ** i - 1
        r = int(s.to_integral_value(rounding=ROUND_CEILING)) + 1
        n = (r ** 3 + 7) // 8 * 8
        if n <= T:
            first += 1
        else:
            n = T
        res.append(r)
        sizes.append(n)
    return dict(L=L, T=T, N_min=nmin, N_max=nmax, b=float(b), resolutions=res, first_hash_level=first,
                dense_sizes=sizes[:first], n_rows=sum(sizes))


cases = [level_table(16, 2 ** 19, 16, 2048), level_table(16, 2 ** 22, 16, 2048), level_table(5, 2 ** 17, 16, 512),
         level_table(8, 2 ** 12, 4, 64), level_table(4, 2 ** 6, 2, 6)]
quoted = [(1.381913, 5, [4096, 12168, 29792, 79512, 205384], 6098120), (None, 8, None, 39474024),
          (2.378414, 2, None, 456632)]
for c, (b, fhl, dense, rows) in zip(cases, quoted):
    if b is not None:
        assert abs(c["b"] - b) < 1e-6, (c["b"], b)
    assert c["first_hash_level"] == fhl
    if dense is not None:
        assert c["dense_sizes"] == dense
    assert c["n_rows"] == rows, (c["n_rows"], rows)

out = Path(__file__).resolve().parent / "hashgrid_levels.json"
out.write_text(json.dumps({"source": "tests/golden/make_golden.py", "cases": cases}, indent=1) + "\n")
print("wrote", out)
