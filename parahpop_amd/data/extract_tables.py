"""Write the published constant tables the force model needs as plain data.

The three tables are published physical-model data, not code, and must hold
the same numbers as the reference for the force model to agree with it:

* NRLMSISE-00 coefficient blocks (pt, pd, ps, pdl, ptm, pdm, ptl, pma, sam,
  pavgm) of the NRL model, as held in
  src/common/engine/paraHPOP/model/environment/atmosphere/detail/Nrlmsise00Data.inl
* IERS 2010 Tables 6.5a/b/c frequency-dependent solid-tide corrections, as held
  in src/common/engine/paraHPOP/model/accelerations/hpop_compat/HpopFrequencyTerms.inc
* FES2004 ocean-tide Stokes amplitudes (18 waves, 402 records), as held in
  src/common/engine/paraHPOP/model/accelerations/hpop_compat/Fes2004Terms.inc

Run once in a container that has the reference checkout; the outputs
(``*.txt`` beside this script) are committed and are what the library and the
oracle load at run time.  Usage::

    python parahpop_amd/data/extract_tables.py /root/reference
"""
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
NUM = r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?"


def strip_comments(text):
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return re.sub(r"//[^\n]*", " ", text)


def block(text, name):
    m = re.search(name + r"\s*(\[[^=]*\])\s*=\s*\{", text)
    if not m:
        raise SystemExit("table %s not found" % name)
    depth, i = 1, m.end()
    while depth:
        c = text[i]
        depth += c == "{"
        depth -= c == "}"
        i += 1
    dims = [int(d) for d in re.findall(r"\[(\d+)\]", m.group(1))]
    return dims, text[m.end():i - 1]


def nrlmsise(ref):
    path = os.path.join(ref, "src/common/engine/paraHPOP/model/environment/atmosphere/detail/Nrlmsise00Data.inl")
    text = strip_comments(open(path).read())
    out = ["# NRLMSISE-00 coefficient blocks: 'name dim...' line, then the values row-major"]
    for name in ["pt", "pd", "ps", "pdl", "ptm", "pdm", "ptl", "pma", "sam", "pavgm"]:
        dims, body = block(text, r"\b" + name)
        vals = [float(v) for v in re.findall(NUM, body)]
        count = 1
        for d in dims:
            count *= d
        if len(vals) != count:
            raise SystemExit("%s: expected %d values, found %d" % (name, count, len(vals)))
        out.append(name + " " + " ".join(str(d) for d in dims))
        for k in range(0, len(vals), 5):
            out.append(" ".join(repr(v) for v in vals[k:k + 5]))
    return "\n".join(out) + "\n"


def frequency_terms(ref):
    path = os.path.join(ref, "src/common/engine/paraHPOP/model/accelerations/hpop_compat/HpopFrequencyTerms.inc")
    text = strip_comments(open(path).read())
    out = ["# IERS 2010 Tables 6.5a (order 0), 6.5b (order 1), 6.5c (order 2)",
           "# order  n1 n2 n3 n4 n5 (Delaunay multipliers l l' F D Omega)  in-phase  out-of-phase  (units 1e-12)"]
    for order, name in ((0, "terms20"), (1, "terms21"), (2, "terms22")):
        m = re.search(r"HPOP_FREQUENCY_TABLE\(\s*%s\s*,\s*(\d+)\s*\)\s*=\s*\{" % name, text)
        if not m:
            raise SystemExit("frequency table %s not found" % name)
        end = text.index("};", m.end())
        rows = re.findall(r"\{\s*\{([^}]*)\}\s*,\s*(%s)\s*,\s*(%s)\s*\}" % (NUM, NUM), text[m.end():end])
        if len(rows) != int(m.group(1)):
            raise SystemExit("%s: expected %s rows, found %d" % (name, m.group(1), len(rows)))
        for mult, ip, op in rows:
            mm = [int(v) for v in mult.split(",")]
            out.append("%d %s %r %r" % (order, " ".join(str(v) for v in mm), float(ip), float(op)))
    return "\n".join(out) + "\n"


def fes2004(ref):
    path = os.path.join(ref, "src/common/engine/paraHPOP/model/accelerations/hpop_compat/Fes2004Terms.inc")
    raw = open(path).read()
    names = re.findall(r"//\s*([\d.]+)\s+(\w+)\s*$", raw, flags=re.M)
    text = strip_comments(raw)
    m = re.search(r"FES_WAVES\(\s*waves\s*,\s*(\d+)\s*\)\s*=\s*\{", text)
    end = text.index("};", m.end())
    waves = re.findall(r"\{\s*\{([^}]*)\}\s*,\s*(\d+)\s*,\s*(\d+)\s*\}", text[m.end():end])
    m2 = re.search(r"FES_TERMS\(\s*terms\s*,\s*(\d+)\s*\)\s*=\s*\{", text)
    end2 = text.index("};", m2.end())
    terms = re.findall(r"\{\s*(\d+)\s*,\s*(\d+)\s*,\s*(%s)\s*,\s*(%s)\s*,\s*(%s)\s*,\s*(%s)\s*\}" % (NUM, NUM, NUM, NUM),
                       text[m2.end():end2])
    if len(waves) != int(m.group(1)) or len(terms) != int(m2.group(1)):
        raise SystemExit("FES2004 table sizes do not match their declarations")
    out = ["# FES2004 ocean-tide constituents (dimensionless, file values * 1e-11)",
           "# wave lines: W doodson1..6 first_term end_term darwin_number name",
           "# term lines: T n m Cp Sp Cm Sm"]
    for k, (dood, b, e) in enumerate(waves):
        num, name = names[k] if k < len(names) else ("0", "?")
        out.append("W %s %s %s %s %s" % (" ".join(v.strip() for v in dood.split(",")), b, e, num, name))
    for n, mm, cp, sp, cm, sm in terms:
        out.append("T %s %s %r %r %r %r" % (n, mm, float(cp), float(sp), float(cm), float(sm)))
    return "\n".join(out) + "\n"


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    for fname, fn in (("nrlmsise00_coefficients.txt", nrlmsise),
                      ("iers2010_solid_tide_frequency.txt", frequency_terms),
                      ("fes2004_ocean_tide.txt", fes2004)):
        with open(os.path.join(HERE, fname), "w") as f:
            f.write(fn(ref))
        print("wrote", fname)


if __name__ == "__main__":
    main()
