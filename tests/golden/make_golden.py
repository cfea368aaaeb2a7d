"""Writes the golden vectors in tests/golden/*.npz from the reference itself.

Runs oracle/_ref/ref_harness (the reference's module sht, shtransforms,
plms, fft, horizontal and get_nl compiled from the reference sources by
oracle/Makefile) on seeded synthetic inputs and stores inputs and outputs.
Run in the build container, where the reference tree exists:

    make -C oracle && python tests/golden/make_golden.py
"""
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from magic_amd import workload  # noqa: E402
from oracle import oracle  # noqa: E402

sys.path.insert(0, os.path.join(ROOT, "tests"))
from api_cases import read_header, write_api_input, read_api_output  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))

def make_api_inputs(l_max, minc, n_theta, n_phi, lcuts, stream, real_m0=True):
    ls, ms = workload.lm_index(l_max, None, minc)
    levels = []
    for lev, lcut in enumerate(lcuts):
        coefs = []
        for fld in range(11):
            c = workload.coefficients(l_max, stream, fld, lev, "l+1^-2", potential=False,
                                      minc=minc)
            if not real_m0:
                u = workload.uniform01(workload.stream_seed(stream, 50 + fld, lev), ls.size)
                c = c + 1j * (ms == 0) * (u - 0.5) / (ls + 1.0) ** 2
            coefs.append(c)
        u = workload.uniform01(workload.stream_seed(stream, 99, lev), 3 * n_phi * n_theta)
        grids = (2.0 * u - 1.0).reshape(3, n_phi, n_theta)
        levels.append(dict(lcut=lcut, r=0.3, r_icb=0.35, or2=1.0 / 0.81, coefs=coefs,
                           grids=grids))
    return levels


def api_case(tag, l_max, minc, lcuts, stream, real_m0=True):
    n_theta, n_phi = workload.grid_size(l_max, 20, minc)
    levels = make_api_inputs(l_max, minc, n_theta, n_phi, lcuts, stream, real_m0)
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        write_api_input(fin, levels)
        oracle.run_ref("api", l_max, minc, 20, fin, fout, nlev=len(levels), threads=4)
        hd, res = read_api_output(fout, len(levels))
    assert hd["n_theta"] == n_theta and hd["n_phi"] == n_phi, hd
    data = dict(l_max=l_max, minc=minc, n_theta=n_theta, n_phi=n_phi, m_max=hd["m_max"],
                lcut=np.array(lcuts), r=0.3, r_icb=0.35, or2=1.0 / 0.81)
    for lev, (L, R) in enumerate(zip(levels, res)):
        data["in_coefs_%d" % lev] = np.stack(L["coefs"])
        data["in_grids_%d" % lev] = L["grids"]
        for k, v in R.items():
            data["out_%s_%d" % (k, lev)] = v
    path = os.path.join(OUT, "api_%s.npz" % tag)
    np.savez_compressed(path, **data)
    print("wrote", path, os.path.getsize(path))


def tables_case(tag, l_max, minc):
    n_theta, n_phi = workload.grid_size(l_max, 20, minc)
    with tempfile.TemporaryDirectory() as td:
        fout = os.path.join(td, "t.bin")
        oracle.run_ref("tables", l_max, minc, 20, "/dev/null", fout, threads=1)
        with open(fout, "rb") as f:
            hd = read_header(f)
            nt, lm = hd["n_theta"], hd["lm_max"]
            theta_ord = np.fromfile(f, "<f8", nt)
            gauss = np.fromfile(f, "<f8", nt)
            osin2 = np.fromfile(f, "<f8", nt)
            cost = np.fromfile(f, "<f8", nt)
            sint = np.fromfile(f, "<f8", nt)
            tabs = [np.fromfile(f, "<f8", lm * (nt // 2)).reshape(nt // 2, lm) for _ in range(4)]
            lm2l = np.fromfile(f, "<f8", lm)
            lm2m = np.fromfile(f, "<f8", lm)
            assert f.read() == b""
    path = os.path.join(OUT, "tables_%s.npz" % tag)
    np.savez_compressed(path, l_max=l_max, minc=minc, n_theta=n_theta, n_phi=n_phi,
                        theta_ord=theta_ord, gauss=gauss, osin2=osin2, cos_theta=cost,
                        sin_theta=sint, plm=tabs[0], dplm=tabs[1], wplm=tabs[2],
                        wdplm=tabs[3], lm2l=lm2l.astype(np.int32), lm2m=lm2m.astype(np.int32))
    print("wrote", path, os.path.getsize(path))


def nl_case(tag, l_max):
    n_theta, n_phi = workload.grid_size(l_max, 20, 1)
    u = workload.uniform01(workload.stream_seed(7, 0, 0), 13 * n_phi * n_theta)
    grids = (2.0 * u - 1.0).reshape(13, n_phi, n_theta)
    rad = np.array([1.0 / 0.81, 1.0 / 0.6561, 1.25, 0.8])  # or2, or4, orho1, LFfac
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            rad.astype("<f8").tofile(f)
            grids.astype("<f8").tofile(f)
        oracle.run_ref("nl", l_max, 1, 20, fin, fout, threads=2)
        with open(fout, "rb") as f:
            read_header(f)
            out = np.fromfile(f, "<f8", 12 * n_phi * n_theta).reshape(12, n_phi, n_theta)
            assert f.read() == b""
    path = os.path.join(OUT, "nl_%s.npz" % tag)
    np.savez_compressed(path, l_max=l_max, n_theta=n_theta, n_phi=n_phi, radial=rad,
                        in_grids=grids, out=out)
    print("wrote", path, os.path.getsize(path))


def nlstep_case(tag, l_max, nlev, write=True):
    """The nonlinear-term step (ref_harness nlstep mode) on nl_reference's inputs."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from nl_reference import make_inputs, radial
    n_theta, n_phi = workload.grid_size(l_max)
    inputs = make_inputs(l_max, nlev)
    or2, or4, orho1, lff = radial(nlev)
    lcut = [l_max] * nlev
    lcut[-1] = l_max - 3
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            for i in range(nlev):
                np.array([lcut[i]], dtype="<i4").tofile(f)
                np.array([or2[i], or4[i], orho1[i], lff], dtype="<f8").tofile(f)
                for k in range(11):
                    np.ascontiguousarray(inputs[k][i], dtype="<c16").tofile(f)
        oracle.run_ref("nlstep", l_max, 1, 20, fin, fout, nlev=nlev, threads=4)
        with open(fout, "rb") as f:
            hd = read_header(f)
            lm = hd["lm_max"]
            G = np.zeros((nlev, 18, n_phi, n_theta))
            O = np.zeros((10, nlev, lm), dtype=np.complex128)
            for i in range(nlev):
                G[i] = np.fromfile(f, "<f8", 18 * n_phi * n_theta).reshape(18, n_phi, n_theta)
                for k in range(10):
                    O[k, i] = np.fromfile(f, "<c16", lm)
            assert f.read() == b""
    data = dict(l_max=l_max, nlev=nlev, n_theta=n_theta, n_phi=n_phi, lcut=np.array(lcut),
                grids=G, outs=O)
    if write:
        path = os.path.join(OUT, "nlstep_%s.npz" % tag)
        np.savez_compressed(path, **data)
        print("wrote", path, os.path.getsize(path))
    return data


TD_IN = ("AdvrLM", "AdvtLM", "AdvpLM", "VStLM", "VxBrLM", "VxBtLM", "VxBpLM", "w", "dw", "z")
TD_OUT = ("dwdt", "dzdt", "dpdt", "dsdt", "dbdt", "djdt", "dVxBhLM")


def td_inputs(l_max, nlev, stream=21):
    """Seeded inputs of the lm-space assembly: the seven nonlinear outputs it
    reads and the fields w, dw/dr, z at each level (complex, (l+1)^-2)."""
    return np.stack([np.stack([workload.coefficients(l_max, stream, k, lev, "l+1^-2", potential=False)
                               for lev in range(nlev)]) for k in range(len(TD_IN))])


def td_radial(nlev):
    r = np.linspace(0.6, 1.4, nlev)
    return 1.0 / r ** 2, 1.0 / r ** 4, 1.0 / (1.0 + 0.1 * r), 250.0, 0.7


def td_case(tag, l_max, nlev, write=True):
    """The reference's get_td (ref_harness td mode): dwdt, dzdt, dpdt, dsdt,
    dbdt, djdt, dVxBhLM of interior levels, one with a lower l_R."""
    X = td_inputs(l_max, nlev)
    or2, or4, orho1, corfac, epsc = td_radial(nlev)
    lcut = [l_max] * nlev
    lcut[-1] = l_max - 4
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            for i in range(nlev):
                np.array([lcut[i]], dtype="<i4").tofile(f)
                np.array([or2[i], or4[i], orho1[i], corfac, epsc], dtype="<f8").tofile(f)
                for k in range(len(TD_IN)):
                    np.ascontiguousarray(X[k, i], dtype="<c16").tofile(f)
        oracle.run_ref("td", l_max, 1, 20, fin, fout, nlev=nlev, threads=1)
        with open(fout, "rb") as f:
            hd = read_header(f)
            lm = hd["lm_max"]
            O = np.zeros((len(TD_OUT), nlev, lm), dtype=np.complex128)
            for i in range(nlev):
                for k in range(len(TD_OUT)):
                    O[k, i] = np.fromfile(f, "<c16", lm)
            assert f.read() == b""
    data = dict(l_max=l_max, nlev=nlev, lcut=np.array(lcut), inputs=X, outs=O,
                or2=or2, or4=or4, orho1=orho1, corfac=corfac, epsc=epsc)
    if write:
        path = os.path.join(OUT, "td_%s.npz" % tag)
        np.savez_compressed(path, **data)
        print("wrote", path, os.path.getsize(path))
    return data


def courant_case(tag, l_max, nlev, write=True):
    """The reference's Courant limits (src/courant.f90, magnetic branch with
    Alfven damping) on the grids of the nonlinear step (ref_harness courant,
    nl_reference's inputs)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from nl_reference import make_inputs, radial
    inputs = make_inputs(l_max, nlev)
    or2, or4, orho1, lff = radial(nlev)
    lcut = [l_max] * nlev
    lcut[-1] = l_max - 3
    courfac, alffac, opm = 2.5, 1.0, 0.2
    delxr2 = 1.0e-3 * (1.0 + np.arange(nlev))
    delxh2 = 2.0e-2 / (1.0 + np.arange(nlev))
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            for i in range(nlev):
                np.array([lcut[i]], dtype="<i4").tofile(f)
                np.array([or2[i], or4[i], orho1[i], lff, courfac, alffac, opm, delxr2[i], delxh2[i]],
                         dtype="<f8").tofile(f)
                for k in range(11):
                    np.ascontiguousarray(inputs[k][i], dtype="<c16").tofile(f)
        oracle.run_ref("courant", l_max, 1, 20, fin, fout, nlev=nlev, threads=1)
        with open(fout, "rb") as f:
            read_header(f)
            dt = np.fromfile(f, "<f8", 2 * nlev).reshape(nlev, 2)
            assert f.read() == b""
    data = dict(l_max=l_max, nlev=nlev, lcut=np.array(lcut), courfac=courfac, alffac=alffac,
                opm=opm, delxr2=delxr2, delxh2=delxh2, dtrkc=dt[:, 0], dthkc=dt[:, 1])
    if write:
        path = os.path.join(OUT, "courant_%s.npz" % tag)
        np.savez_compressed(path, **data)
        print("wrote", path, os.path.getsize(path))
    return data


def brv_case(tag, l_max, minc, stream, write=True):
    """The reference's boundary products of a conducting mantle and inner
    core (src/nonlinear_bcs.f90 get_br_v_bcs, ref_harness brv): two levels,
    the outer boundary with omega_ma and the inner one with omega_ic, on
    seeded grids uniform in [-1, 1]; and on those products the nonlinear
    boundary terms of get_b_nl_bcs for 'CMB' (b_nl_bc, aj_nl_bc) and 'ICB'
    (aj_nl_bc) with conductance_ma, prmag, sigma_ratio per level."""
    n_theta, n_phi = workload.grid_size(l_max, 20, minc)
    nlev = 2
    omega = np.array([0.7, -1.3])
    or2 = np.array([1.0 / 1.538 ** 2, 1.0 / 0.538 ** 2])
    orho = np.array([1.0, 0.85])
    cond = np.array([0.3, 1.7])
    prmag = np.array([5.0, 0.45])
    sigma = np.array([1.0, 0.61])
    grids = []
    for lev in range(nlev):
        u = workload.uniform01(workload.stream_seed(stream, 77, lev), 3 * n_phi * n_theta)
        grids.append((2.0 * u - 1.0).reshape(3, n_phi, n_theta))
    grids = np.stack(grids)
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            for i in range(nlev):
                np.array([omega[i], or2[i], orho[i], cond[i], prmag[i], sigma[i]], dtype="<f8").tofile(f)
                np.ascontiguousarray(grids[i], dtype="<f8").tofile(f)
        oracle.run_ref("brv", l_max, minc, 20, fin, fout, nlev=nlev, threads=1)
        with open(fout, "rb") as f:
            hdr = read_header(f)
            lm_max = hdr["lm_max"]
            out = np.fromfile(f, "<c16", 5 * nlev * lm_max).reshape(nlev, 5, lm_max)
            assert f.read() == b""
    data = dict(l_max=l_max, minc=minc, n_theta=n_theta, n_phi=n_phi, omega=omega, or2=or2,
                orho=orho, grids=grids, br_vt_lm=out[:, 0], br_vp_lm=out[:, 1],
                conductance_ma=cond, prmag=prmag, sigma_ratio=sigma, b_nl_cmb=out[:, 2],
                aj_nl_cmb=out[:, 3], aj_nl_icb=out[:, 4])
    if write:
        path = os.path.join(OUT, "brv_%s.npz" % tag)
        np.savez_compressed(path, **data)
        print("wrote", path, os.path.getsize(path))
    return data


def rloop_case(tag, write=True):
    """The bulk levels of the reference's radial loop (ref_harness rloop:
    nl_level, courant, get_td per level) on the interior levels of the
    boussBenchSat restart state (tests/restart_state.py)."""
    import restart_state
    h, ins, P = restart_state.load(np.arange(1, 32))
    nlev = ins[0].shape[0]
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            for i in range(nlev):
                np.array([P["lcut"][i]], dtype="<i4").tofile(f)
                np.array([P["or2"][i], P["or4"][i], P["orho1"][i], P["LFfac"], P["CorFac"],
                          P["epsc"][i], P["courfac"], P["alffac"], P["opm"], P["delxr2"][i],
                          P["delxh2"][i]], dtype="<f8").tofile(f)
                for k in range(11):
                    np.ascontiguousarray(ins[k][i], dtype="<c16").tofile(f)
        oracle.run_ref("rloop", h["l_max"], h["minc"], h["nalias"], fin, fout, nlev=nlev,
                       threads=1)
        lm = h["lm_max"]
        td_out, dt = [], []
        with open(fout, "rb") as f:
            read_header(f)
            for i in range(nlev):
                td_out.append(np.fromfile(f, "<c16", 7 * lm).reshape(7, lm))
                dt.append(np.fromfile(f, "<f8", 2))
            assert f.read() == b""
    td_out = np.stack(td_out, axis=1)  # (7, nlev, lm)
    dt = np.array(dt)
    data = dict(levels=np.arange(1, 32), td=td_out, dtrkc=dt[:, 0], dthkc=dt[:, 1])
    if write:
        path = os.path.join(OUT, "rloop_%s.npz" % tag)
        np.savez_compressed(path, **data)
        print("wrote", path, os.path.getsize(path))
    return data


def main():
    if not oracle.have_ref():
        sys.exit("oracle/_ref/ref_harness is not built (make -C oracle)")
    tables_case("L16_m1", 16, 1)
    tables_case("L64_m4", 64, 4)
    api_case("L16_m1", 16, 1, [16, 11, 0], stream=1, real_m0=True)
    api_case("L16_m1_cplxm0", 16, 1, [16], stream=11, real_m0=False)
    api_case("L16_m4", 16, 4, [16, 9], stream=12)
    api_case("L64_m1", 64, 1, [64], stream=1)
    api_case("L64_m4", 64, 4, [64, 40], stream=13)
    nl_case("L16", 16)
    nlstep_case("L16", 16, 3)
    td_case("L16", 16, 3)
    td_case("L64", 64, 2)
    courant_case("L16", 16, 3)
    brv_case("L16_m1", 16, 1, stream=21)
    brv_case("L64_m4", 64, 4, stream=22)
    rloop_case("boussBenchSat")


if __name__ == "__main__":
    main()
