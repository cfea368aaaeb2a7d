"""Writes make rules for compiling the reference's Fortran sources in
dependency order (test infrastructure: the full reference build under
oracle/_ref/full, see oracle/Makefile).

    python refdeps.py REFSRC OBJDIR > deps.mk

Each object depends on the objects of the modules its source `use`s, so
that `make -j` compiles independent modules in parallel.  The build takes
the reference's serial, in-tree library choices (src/Makefile:250-300 with
USE_MPI=no, USE_FFTLIB=JW, USE_DCTLIB=JW, USE_LAPACKLIB=JW, USE_SHTNS=no):
fft.f90, cosine_transform_odd.f90, algebra.f90 and sht_native.f90, not their
FFTW / LAPACK / SHTns alternatives.
"""
import os
import re
import sys

EXCLUDED = {"fft_fftw.f90", "dct_fftw.f90", "algebra_lapack.f90", "shtns.f90"}


def scan(src):
    files = sorted(f for f in os.listdir(src) if f.endswith(".f90") and f not in EXCLUDED)
    mod2file, uses = {}, {}
    for f in files:
        text = open(os.path.join(src, f), errors="replace").read()
        for m in re.finditer(r"^\s*module\s+(\w+)\s*$", text, re.M | re.I):
            mod2file[m.group(1).lower()] = f
        uses[f] = sorted(set(u.lower() for u in re.findall(r"^\s*use\s+(\w+)", text, re.M | re.I)))
    deps = {}
    for f in files:
        d = []
        for u in uses[f]:
            g = mod2file.get(u)
            if g and g != f and g not in d:
                d.append(g)
        deps[f] = d
    return files, deps


def main():
    src, obj = sys.argv[1], sys.argv[2]
    files, deps = scan(src)
    stems = [f[:-4] for f in files]
    print("REF_FULL_OBJS := " + " ".join("%s/%s.o" % (obj, s) for s in stems))
    for f in files:
        s = f[:-4]
        dl = " ".join("%s/%s.o" % (obj, d[:-4]) for d in deps[f])
        print("%s/%s.o: %s/%s %s" % (obj, s, src, f, dl))
        print("\t$(REF_FC) $(FULLFLAGS) -module-dir %s -I%s -c %s/%s -o $@" % (obj, obj, src, f))


if __name__ == "__main__":
    main()
