"""Extract the STO -> Gaussian fit tables (ng = 1..30) as plain data.

The reference keeps the least-squares fits of exp(-r) by
sum_i c_i exp(-g_i r^2) as 930 assignment statements in
``src/read_fit_SMILES.F90:4-1398`` (``c_fit_exp(i,ng)``, ``g_fit_exp(i,ng)``).
They are numbers, not code: this script reads that file as text and writes
``sto_fit_tables.txt`` beside it, one line per (ng, i) with the decimal
strings kept verbatim (Fortran ``D`` exponent turned into ``E``) so the
values parse to the same doubles.

Run (only where the reference tree is mounted):
    python slater_stoch_amd/data/extract_fit_tables.py /root/reference/src/read_fit_SMILES.F90
"""
import re
import sys
from pathlib import Path

PAT = re.compile(r"^\s*([cg])_fit_exp\(\s*(\d+)\s*,\s*(\d+)\s*\)\s*=\s*([-+0-9.DdEe]+)\s*$")


def extract(src: Path) -> dict:
    table = {}
    for line in src.read_text().splitlines():
        m = PAT.match(line)
        if not m:
            continue
        kind, i, ng, val = m.group(1), int(m.group(2)), int(m.group(3)), m.group(4)
        table.setdefault((ng, i), {})[kind] = val.replace("D", "E").replace("d", "E")
    return table


def main():
    src = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src/read_fit_SMILES.F90")
    table = extract(src)
    ngs = sorted({ng for ng, _ in table})
    assert ngs == list(range(1, 31)), ngs
    for ng in ngs:
        assert sorted(i for n, i in table if n == ng) == list(range(1, ng + 1)), ng
    out = Path(__file__).with_name("sto_fit_tables.txt")
    with out.open("w") as f:
        f.write("# ng i c_fit g_fit : exp(-r) ~ sum_i c_fit(i,ng) exp(-g_fit(i,ng) r^2)\n")
        f.write("# values from src/read_fit_SMILES.F90 of scemama/slater_stoch (data only)\n")
        for ng, i in sorted(table):
            e = table[(ng, i)]
            f.write(f"{ng:2d} {i:2d} {e['c']} {e['g']}\n")
    print(f"wrote {out} ({len(table)} entries)")


if __name__ == "__main__":
    main()
