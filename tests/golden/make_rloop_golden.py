"""Writes tests/golden/rloop_<case>.npz: one call of the reference's own
radial loop (rIter_single_t%radialLoop, src/rIter.f90:94-444) under the
physics switches of a run, inputs and outputs, as the reference program
itself computes them.

Each case runs oracle/_ref/magic_dump (the reference's magic.exe compiled
from its sources by oracle/Makefile, with the radial-loop module replaced by
the dumping wrapper oracle/ref_rloop_dump.f90) on an input namelist below
(from a restart file of the reference's samples, or from the reference's
own initial fields after a few steps), and stores the dump of one
radialLoopG call for a subset of radial levels (both boundaries and a spread
of interior levels).  Run in the build container:

    make -C oracle full && python tests/golden/make_rloop_golden.py [case ...]
"""
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.dumpfile import read_dump  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
MAGIC_DUMP = os.path.join(ROOT, "oracle", "_ref", "magic_dump")

OUTPUT_CONTROL = """&output_control
 n_log_step  =1000,
 n_graphs    =0,
 n_rsts      =0,
 n_stores    =0,
 n_specs     =0,
 runid       ="golden",
 l_movie     =.false.,
 l_RMS       =.false.,
 l_power     =.false.,
 l_average   =.false.,
/
"""


def nml(grid, control, phys, start, mantle="nRotMa=0", inner="sigma_ratio=0.d0, nRotIC=0"):
    return ("&grid\n%s\n/\n&control\n%s\n/\n&phys_param\n%s\n/\n&start_field\n%s\n/\n%s"
            "&mantle\n %s\n/\n&inner_core\n %s\n/\n" % (grid, control, phys, start, OUTPUT_CONTROL,
                                                         mantle, inner))


BOUSS_GRID = " n_r_max=33, n_cheb_max=31, n_phi_tot=192, n_r_ic_max=17, n_cheb_ic_max=15, minc=4,"
BOUSS_PHYS = (" ra=1.1D5, ek=1.0D-3, pr=1.0D0, prmag=5.0D0, radratio=0.35D0, ktops=1, kbots=1,"
              " ktopv=2, kbotv=2, kbotb=3,")
SMALL_GRID = " n_r_max=17, n_cheb_max=15, n_phi_tot=64, n_r_ic_max=9, n_cheb_ic_max=7, minc=1,"

# case: (namelist, restart file under tests/golden or None, dump call number)
CASES = {
    # samples/boussBenchSat: Boussinesq MHD, curl advection, rigid walls,
    # conducting rotating inner core
    "bouss_mhd": (nml(BOUSS_GRID,
                      " mode=0, tag='g', n_time_steps=2, dtmax=2.0D-4, alpha=0.6D0,"
                      " time_scheme='BPR353',",
                      BOUSS_PHYS,
                      " l_start_file=.true., start_file='checkpoint.start', init_b1=0, init_s1=0,",
                      inner="sigma_ratio=1.d0, nRotIC=1"),
                  "boussBenchSat_checkpoint_end.start", 2),
    # the same state with u.grad(u) advection (the reference's double-curl
    # form, polo_flow_eq='DC', stops with a segmentation fault in this serial
    # amdflang build before its first step, so it has no golden here)
    "bouss_nocurl": (nml(BOUSS_GRID,
                            " mode=0, tag='g', n_time_steps=2, dtmax=2.0D-4, alpha=0.6D0,"
                            " time_scheme='CNAB2', l_adv_curl=.false.,",
                            BOUSS_PHYS,
                            " l_start_file=.true., start_file='checkpoint.start', init_b1=0,"
                            " init_s1=0,",
                            inner="sigma_ratio=1.d0, nRotIC=1"),
                        "boussBenchSat_checkpoint_end.start", 1),
    # samples/doubleDiffusion: non-magnetic convection with composition
    "doublediff": (nml(BOUSS_GRID,
                       " mode=1, tag='g', n_time_steps=2, dtmax=3.0D-4, alpha=0.6D0,"
                       " time_scheme='BPR353',",
                       " ra=4.8d4, raxi=1.2d5, ek=1.0d-3, pr=0.3d0, sc=3.0d0, radratio=0.35d0,"
                       " ktops=1, kbots=1, ktopxi=1, kbotxi=1, ktopv=2, kbotv=2,",
                       " l_start_file=.true., start_file='checkpoint.start', init_s1=0,"
                       " init_xi1=0,"),
                   "doubleDiffusion_checkpoint_end.start", 3),
    # anelastic dynamo (the physics of samples/hydro_bench_anel with a
    # magnetic field and the centrifugal term): u.grad(u), viscous and Ohmic
    # heating; stress-free walls
    "anel_mhd": (nml(SMALL_GRID,
                     " mode=0, tag='g', n_time_steps=20, dtmax=1.0D-4, alpha=0.6D0,"
                     " courfac=2.5D0, alffac=1.0D0,",
                     " ra=1.48638035D5, ek=1.0D-3, pr=1.0D0, prmag=5.0D0, strat=5.D0,"
                     " polind=2.0D0, radratio=0.35D0, g0=0.D0, g1=0.D0, g2=1.D0, ktops=1,"
                     " kbots=1, ktopv=1, kbotv=1, dilution_fac=1.0D-3,",
                     " l_start_file=.false., init_b1=3, amp_b1=5, init_s1=1919, amp_s1=0.01,"
                     " init_v1=3, amp_v1=3.0,"),
                 None, 20),
    # precession (samples/precession physics), non-magnetic Boussinesq
    "precession": (nml(SMALL_GRID,
                       " mode=1, tag='g', n_time_steps=20, dtmax=1.0D-3, alpha=0.6D0,",
                       " ra=0.0D0, ek=1.0D-3, pr=1.0D0, po=-1.0d-2, prec_angle=23.5d0,"
                       " radratio=0.35D0, ktopv=2, kbotv=2,",
                       " l_start_file=.false., init_s1=0, init_v1=3, amp_v1=0.5,"),
                   None, 20),
    # phase field (samples/phase_field physics, Chebyshev grid)
    "phase": (nml(SMALL_GRID,
                  " mode=1, tag='g', n_time_steps=20, dtmax=1.0D-4, alpha=0.6D0, courfac=2.5D0,"
                  " alffac=1.0D0,",
                  " ra=2.0D5, ek=1.0D-3, pr=1.0D0, radratio=0.35D0, stef=1.0d0, tmelt=0.11d0,"
                  " phaseDiffFac=1.0d0, penaltyFac=0.5d0, epsPhase=0.03d0, ktops=1, kbots=1,"
                  " ktopv=2, kbotv=2, ktopphi=2, kbotphi=2,",
                  " l_start_file=.false., init_s1=0404, amp_s1=0.1d0, init_phi=1, init_v1=3,"
                  " amp_v1=0.5,"),
              None, 20),
}

# radial levels kept (1-based nR): boundaries and a spread of interior ones
def keep_levels(n_r):
    return sorted(set([1, 2, 3, n_r // 4, n_r // 2, (3 * n_r) // 4, n_r - 2, n_r - 1, n_r]))


PER_LEVEL = ("r", "or1", "or2", "or3", "or4", "orho1", "orho2", "rho0", "beta", "otemp1", "temp0",
             "visc", "lambda", "epscProf", "l_R", "delxr2", "delxh2", "dtrkc", "dthkc")
FIELDS = ("w", "dw", "ddw", "z", "dz", "p", "s", "ds", "xi", "phi", "b", "db", "ddb", "aj", "dj",
          "dwdt", "dzdt", "dpdt", "dsdt", "dxidt", "dphidt", "dVSrLM", "dVXirLM", "dbdt", "djdt",
          "dVxVhLM", "dVxBhLM")


def run_case(name):
    text, start, call = CASES[name]
    with tempfile.TemporaryDirectory() as td:
        with open(os.path.join(td, "input.nml"), "w") as f:
            f.write(text)
        if start:
            shutil.copy(os.path.join(OUT, start), os.path.join(td, "checkpoint.start"))
        env = dict(os.environ, MHX_DUMP_CALL=str(call), MHX_DUMP_FILE=os.path.join(td, "dump.bin"),
                   OMP_NUM_THREADS="4", OMP_STACKSIZE="1G")
        r = subprocess.run(["bash", "-c", "ulimit -s unlimited; %s input.nml" % MAGIC_DUMP], cwd=td,
                           env=env, capture_output=True, text=True)
        if r.returncode != 0 or not os.path.exists(os.path.join(td, "dump.bin")):
            sys.stderr.write(r.stdout[-3000:] + r.stderr[-3000:])
            raise RuntimeError("reference run failed for case " + name)
        d = read_dump(os.path.join(td, "dump.bin"))
    n_r = int(d["n_r_max"])
    keep = np.array(keep_levels(n_r))
    out = {"levels": keep.astype(np.int32), "call": np.int32(call)}
    for k, v in d.items():
        if k in PER_LEVEL:
            out[k] = v[keep - 1]
        elif k in FIELDS:
            out[k] = v[keep - 1]
        else:
            out[k] = v
    path = os.path.join(OUT, "rloop_%s.npz" % name)
    np.savez_compressed(path, **out)
    print("wrote %s (%d bytes): l_max=%d minc=%d n_r=%d levels %s" % (
        path, os.path.getsize(path), int(d["l_max"]), int(d["minc"]), n_r, list(keep)))


if __name__ == "__main__":
    names = sys.argv[1:] or list(CASES)
    for n in names:
        run_case(n)
