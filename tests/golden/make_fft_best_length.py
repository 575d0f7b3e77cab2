"""Writes fft_best_length.json: the three length tables of
WireCell::fft_best_length, read from the reference's source text
(util/src/FFTBestLength.cxx:12-58) by parsing its live `edges +=`
statements (commented-out tables are skipped), plus the expected
results at every table boundary (each entry, one below and one above,
and the largest entries + 1) computed from those parsed tables.

The parsed tables are independent of the copies typed into the oracle
and the host library, so a mistyped entry in either fails the test that
reads this file.  Run from anywhere:

    python tests/golden/make_fft_best_length.py
"""
import json
import os
import re

SRC = "/root/reference/util/src/FFTBestLength.cxx"


def parse_tables(text):
    # strip // comments, then take the numbers of each `edges += ...;`
    live = "\n".join(line.split("//", 1)[0] for line in text.splitlines())
    tabs = []
    for m in re.finditer(r"edges\s*\+=\s*([0-9,\s]+);", live):
        tabs.append([int(x) for x in m.group(1).replace("\n", " ").split(",") if x.strip()])
    if len(tabs) != 3:
        raise SystemExit(f"expected 3 live tables, found {len(tabs)}")
    return {"any": tabs[0], "even": tabs[1], "odd": tabs[2]}


def best(tab, n):
    if n <= tab[-1]:
        for e in tab:
            if e >= n:
                return e
    return n


def main():
    with open(SRC) as f:
        tabs = parse_tables(f.read())
    cases = []
    for name, keep in (("any", 0), ("even", 1), ("odd", 1)):
        tab = tabs[name]
        ns = set()
        for e in tab:
            ns.update((e - 1, e, e + 1))
        ns.update((1, 2, 1000, 6000, tab[-1] + 1, 20000))
        for n in sorted(ns):
            if n < 1:
                continue
            if keep and (n % 2 == 0) != (name == "even"):
                continue
            cases.append([n, keep, best(tab, n)])
    out = {"source": "util/src/FFTBestLength.cxx:12-58", "tables": tabs, "cases": cases}
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "fft_best_length.json"), "w") as f:
        json.dump(out, f, indent=0)


if __name__ == "__main__":
    main()
