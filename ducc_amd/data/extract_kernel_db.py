"""Regenerate ``kernel_db.csv`` from ducc0's tabulated ES-kernel database.

The table (support W, oversampling factor, tabulated error, beta, e0, ndim,
single precision) is the output of ducc0's kernel-design scan and is data, not
code: the parameter choice of the plain path must match ducc0's exactly, so the
same rows are used.  Source: src/ducc0/math/gridding_kernel.cc:78-1537.

Usage: python ducc_amd/data/extract_kernel_db.py [path/to/gridding_kernel.cc]
"""
import os
import re
import sys

ROW = re.compile(
    r"^\{\s*(\d+),\s*([0-9.]+),\s*([0-9.eE+-]+),\s*([0-9.eE+-]+),"
    r"\s*([0-9.eE+-]+),\s*(\d),\s*(true|false)\}")


def extract(src):
    rows = []
    with open(src) as f:
        inside = False
        for line in f:
            if "KernelDB {" in line:
                inside = True
                continue
            if inside:
                m = ROW.match(line.strip())
                if m:
                    W, of, eps, beta, e0, nd, sp = m.groups()
                    rows.append((int(W), of, eps, beta, e0, int(nd),
                                 1 if sp == "true" else 0))
                elif line.startswith("}"):
                    break
    return rows


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else \
        "/root/reference/src/ducc0/math/gridding_kernel.cc"
    rows = extract(src)
    assert len(rows) == 1425, len(rows)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                       "kernel_db.csv")
    with open(out, "w") as f:
        f.write("# W,ofactor,epsilon,beta,e0,ndim,singleprec\n")
        for r in rows:
            f.write(",".join(str(x) for x in r) + "\n")
    print(f"wrote {len(rows)} rows to {out}")


if __name__ == "__main__":
    main()
