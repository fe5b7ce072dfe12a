"""Write golden vectors produced by the reference's own Fortran routines.

Runs in the development container only (it needs the reference sources, built by
oracle/build_ref.sh into oracle/_ref/).  The output, tests/golden/ref_small.npz, holds
inputs and the reference routines' outputs for one small channel:

* y_grid_canopy metrics, def_k wavenumbers, gridweighting        (y_grid_canopy.f90, start.f90)
* boundary_canopies point lists                                    (geom_sym_canopy.f90)
* LU_build + LU_dec factors of the u and v matrices, axz           (tridLU_3D_IB.f90:427-640)
* LU_build alone: the unfactored rows of those matrices            (tridLU_3D_IB.f90:427-493)
* LU_buildP + LU_decP factors DG                                   (tridLU_3D_IB.f90:497-684)
* RHS0_u1/u2/u3 outputs du and Lu on a seeded state               (littleharsh.f90:614-731)
* solveP + v_corr on a seeded state                                (littleharsh.f90:557-937)
* flowrate_corr on the (0,0) column                                (littleharsh.f90:453-485)
* buildsp / buildspUV                                              (spectra.f90:18-56)
* stats and the conditional folds conddown1/2, condup1/2 on physical planes made by
  rft/cft from the seeded state                                    (stats.f90:1-253)
* the files the reference writes: y_grid_canopy.dat, boundary_*.dat, stats_*.dat
  (write_stats) and the spectra files (recvwrspec), as raw bytes
* FFTPACK rft/cft forward and inverse on seeded data               (rft_buff.f, cft_buff.f)
* a physical <-> spectral plane transform built from rft/cft with the 3/2 zero padding

Usage: python tests/golden/make_golden.py
"""
from __future__ import annotations

import ctypes
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REFDIR = os.path.join(ROOT, "oracle", "_ref")

# the small channel of the fixtures (nx != nz on purpose)
CFG = dict(nx=16, nz=8, ny_total=19, dsty=3, posth=0.2, dyq=3.0, re=180.0, alp=1.0, bet=2.0,
           mpgx=-1.0, ntilex=4, ntilez=2, dstx=2, dstz=3)


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def plane_to_phys(fft, spec, nx, nz):
    """(nz, mx) spectral plane -> (pz, px) physical plane: 3/2 zero padding, cft inverse
    along z, rft inverse along x (rft_buff.f:77-84, cft_buff.f:48-57)."""
    px, pz, mx = 3 * nx // 2, 3 * nz // 2, nx // 2 + 1
    wsx = np.zeros(2 * px + 18)
    fft.rfti_(ctypes.byref(ctypes.c_int(px)), _dp(wsx))
    wsz = np.zeros(10 * pz + 19)
    fft.cfti_(ctypes.byref(ctypes.c_int(pz)), _dp(wsz))
    h = nz // 2
    pad = np.zeros((pz, mx), complex)
    pad[:h] = spec[:h]
    pad[pz - h:] = spec[h:]
    pad = np.ascontiguousarray(pad)
    fft.cft_(_dp(pad), ctypes.byref(ctypes.c_int(2 * mx)), ctypes.byref(ctypes.c_int(2)),
             ctypes.byref(ctypes.c_int(mx)), ctypes.byref(ctypes.c_int(0)), _dp(wsz))
    bb = np.zeros((pz, px + 2))
    bb[:, 0:2 * mx:2] = pad.real
    bb[:, 1:2 * mx:2] = pad.imag
    fft.rft_(_dp(bb), ctypes.byref(ctypes.c_int(px + 2)), ctypes.byref(ctypes.c_int(pz)),
             ctypes.byref(ctypes.c_int(0)), _dp(wsx))
    return bb[:, :px].copy()


def main():
    if not os.path.exists(os.path.join(REFDIR, "libref.so")):
        subprocess.run(["bash", os.path.join(ROOT, "oracle", "build_ref.sh")], check=True)
    ref = ctypes.CDLL(os.path.join(REFDIR, "libref.so"), mode=os.RTLD_LAZY | os.RTLD_LOCAL)
    fft = ctypes.CDLL(os.path.join(REFDIR, "libfftpack.so"))
    c = CFG
    ny_in = c["ny_total"] - 2 * c["dsty"]
    out = {}
    with tempfile.TemporaryDirectory() as td:
        d = (td + "/").encode()
        ref.ref_setup(ctypes.c_int(c["nx"]), ctypes.c_int(c["nz"]), ctypes.c_int(ny_in),
                      ctypes.c_int(c["dsty"]), ctypes.c_double(c["posth"]),
                      ctypes.c_double(c["dyq"]), ctypes.c_double(c["re"]),
                      ctypes.c_double(c["alp"]), ctypes.c_double(c["bet"]),
                      ctypes.c_double(c["mpgx"]), ctypes.c_int(len(d)), ctypes.c_char_p(d))
        sz = np.zeros(8, np.int32)
        ref.ref_sizes(_ip(sz))
        n40, n43, n30, n33, nn, ncol, b1, b2 = (int(x) for x in sz)
        nyu, nyv, nyp = n43 + 2 - n40, n33 + 2 - n30, n43 - n40
        out["sizes"] = sz
        yu, yv, du, dv = (np.zeros(n) for n in (nyu, nyv, nyu, nyv))
        d2u, d2v, gw = np.zeros(3 * nyu), np.zeros(3 * nyv), np.zeros(2)
        ref.ref_grid(_dp(yu), _dp(yv), _dp(du), _dp(dv), _dp(d2u), _dp(d2v), _dp(gw))
        out.update(yu=yu, yv=yv, dthdyu=du, dthdyv=dv, dyu2i=d2u.reshape(nyu, 3),
                   dyv2i=d2v.reshape(nyv, 3), gw=gw)
        mx = c["nx"] // 2 + 1
        k1x, k2x, k1z, k2z = np.zeros(mx), np.zeros(mx), np.zeros(c["nz"]), np.zeros(c["nz"])
        ref.ref_wavenumbers(_dp(k1x), _dp(k2x), _dp(k1z), _dp(k2z))
        out.update(k1x=k1x, k2x=k2x, k1z=k1z, k2z=k2z)
        ref.ref_canopies(ctypes.c_int(c["ntilex"]), ctypes.c_int(c["ntilez"]),
                         ctypes.c_int(c["dstx"]), ctypes.c_int(c["dstz"]))
        fn = [f for f in os.listdir(td) if f.startswith("boundary_")][0]
        raw = open(os.path.join(td, fn), "rb").read()
        off = 3 * 8 + 4 * 5 * 4
        nl1, nl2 = struct.unpack_from("<2i", raw, off)
        off += 2 * 4 + 8 * 4
        l1 = np.frombuffer(raw, np.int32, 3 * nl1 * 2, off).reshape(2, nl1, 3)
        off += 3 * nl1 * 2 * 4
        l2 = np.frombuffer(raw, np.int32, 3 * nl2 * 2, off).reshape(2, nl2, 3)
        # list_ib(3, n, grid): grid index 0 = vgrid(1), 1 = ugrid(2)
        out.update(list1_v=l1[0].T.copy(), list1_u=l1[1].T.copy(),
                   list2_v=l2[0].T.copy(), list2_u=l2[1].T.copy())
        # velocity LU factors for kRK = 2, dt = 3e-4
        for name, grid, nj in (("u", 2, nyu), ("v", 1, nyv)):
            ay, axz = np.zeros(3 * nj), np.zeros(ncol)
            ref.ref_lu_vel(ctypes.c_int(2), ctypes.c_double(3e-4), ctypes.c_int(grid), _dp(ay), _dp(axz))
            out[f"ay_{name}"] = ay.reshape(nj, 3)
            out[f"axz_{name}"] = axz
            # the same matrix before LU_dec: the system the velocity solve inverts
            ayb, axzb = np.zeros(3 * nj), np.zeros(ncol)
            ref.ref_lu_build(ctypes.c_int(2), ctypes.c_double(3e-4), ctypes.c_int(grid), _dp(ayb), _dp(axzb))
            out[f"ayb_{name}"] = ayb.reshape(nj, 3)
            out[f"axzb_{name}"] = axzb
        dg = np.zeros(3 * nyp * ncol)
        ref.ref_dg(_dp(dg))
        out["dg"] = dg.reshape(ncol, nyp, 3)
        # seeded state
        rng = np.random.default_rng(20261202)

        def cr(*shape):
            return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)

        u1, u2, u3, p = cr(ncol, nyu), cr(ncol, nyv), cr(ncol, nyu), cr(ncol, nyp)
        n1, n2, n3 = cr(ncol, nyu - 2), cr(ncol, nyv - 2), cr(ncol, nyu - 2)
        out.update(in_u1=u1.copy(), in_u2=u2.copy(), in_u3=u3.copy(), in_p=p.copy(),
                   in_nu1=n1.copy(), in_nu2=n2.copy(), in_nu3=n3.copy())

        def cp(a):   # always a fresh copy: the reference routines update arrays in place
            return np.array(a, dtype=np.complex128, order="C", copy=True)

        for krk in (1, 2, 3):
            res = [np.zeros((ncol, n), complex) for n in (nyu, nyv, nyu, nyu, nyv, nyu)]
            args = [cp(a) for a in (u1, u2, u3, p, n1, n2, n3)]
            ref.ref_rhs0(ctypes.c_int(krk), *[_dp(a) for a in args], *[_dp(a) for a in res])
            for nm, a in zip(("du1", "du2", "du3", "lu1", "lu2", "lu3"), res):
                out[f"rhs0_k{krk}_{nm}"] = a
        for krk in (1, 3):
            s1, s2, s3, sp_ = cp(u1), cp(u2), cp(u3), cp(p)
            psi, div = np.zeros((ncol, nyp), complex), np.zeros((ncol, nyp), complex)
            ref.ref_solvep_vcorr(ctypes.c_int(krk), ctypes.c_double(2e-4), _dp(s1), _dp(s2),
                                 _dp(s3), _dp(sp_), _dp(psi), _dp(div))
            out.update({f"sp_k{krk}_u1": s1, f"sp_k{krk}_u2": s2, f"sp_k{krk}_u3": s3,
                        f"sp_k{krk}_p": sp_, f"sp_k{krk}_psi": psi, f"sp_k{krk}_div": div})
        # flowrate on a laminar-like (0,0) column
        col = cp(u1[0].real * 0 + 0.5 * c["re"] * (1 - yu ** 2) + 0.01 * u1[0])
        q = ctypes.c_double()
        ref.ref_flowrate(_dp(col), ctypes.byref(q))
        out["fr_col"] = col.copy()
        out["fr_q"] = np.array([q.value])
        mpg = ctypes.c_double(c["mpgx"])
        g = ctypes.c_double()
        colc = col.copy()
        ref.ref_flowrate_corr(ctypes.c_int(2), ctypes.c_double(2e-4), ctypes.c_double(120.0),
                              ctypes.byref(mpg), _dp(colc), ctypes.byref(g))
        out.update(frc_u=colc, frc_mpgx=np.array([mpg.value]), frc_g=np.array([g.value]))
        # spectra
        for name, grid, a, b in (("u", 2, u1, u1), ("v", 1, u2, u2), ("p", 3, p, p), ("uv", 0, u1, u3)):
            nj = {2: nyu, 1: nyv, 3: nyp, 0: nyu}[grid]
            spv = np.zeros(nj * ncol)
            ref.ref_buildsp(ctypes.c_int(grid), _dp(cp(a)), _dp(cp(b)), _dp(spv))
            out[f"spec_{name}"] = spv.reshape(ncol, nj)
        # statistics: physical planes of u, w, v_itp (u grid), v, w_itp, wx (v grid), p,
        # made from the seeded state with the reference's rft/cft
        ddthetavi = 1.0
        kz = np.repeat(k1z, mx)
        vi = np.zeros_like(u1)
        vi[:, 1:nyu - 1] = 0.5 * (u2[:, :nyu - 2] + u2[:, 1:nyu - 1])
        wi = 0.5 * (u3[:, :nyv] + u3[:, 1:nyv + 1])
        wxs = ddthetavi * dv[None, :] * (u3[:, 1:nyv + 1] - u3[:, :nyv]) - 1j * kz[:, None] * u2
        nxc, nzc = c["nx"], c["nz"]
        planes = {nm: np.stack([plane_to_phys(fft, a[:, j].reshape(nzc, mx), nxc, nzc)
                                for j in range(a.shape[1])])
                  for nm, a in (("u", u1), ("w", u3), ("vi", vi), ("v", u2), ("wi", wi),
                                ("wx", wxs), ("p", p))}
        pxc, pzc = 3 * nxc // 2, 3 * nzc // 2
        dnx, dnz = pxc // c["ntilex"], pzc // c["ntilez"]
        n41 = b1 + 1
        nb = {"u": n41 - n40 + 1, "v": b1 - n30 + 1, "p": n41 - n40 - 1}
        gq = "uuvvuuppuuvvv"
        clen = sum(dnx * dnz * nb[g] for g in gq)
        st = np.zeros(13 * nyu)
        cb, ct = np.zeros(clen), np.zeros(clen)
        ref.ref_stats(ctypes.c_int(c["ntilex"]), ctypes.c_int(c["ntilez"]),
                      *[_dp(np.ascontiguousarray(planes[k])) for k in ("u", "v", "w", "p", "vi", "wi", "wx")],
                      _dp(st), _dp(cb), _dp(ct))
        out.update(stats=st.reshape(13, nyu), cstats_bot=cb, cstats_top=ct)
        # files written by the reference's own writers
        out["file_y_grid"] = np.frombuffer(open(os.path.join(td, "y_grid_canopy.dat"), "rb").read(), np.uint8)
        out["file_boundary"] = np.frombuffer(raw, np.uint8)
        spw = np.zeros(nyu * ncol)
        ref.ref_buildsp(ctypes.c_int(2), _dp(cp(u3)), _dp(cp(u3)), _dp(spw))
        out["spec_w"] = spw.reshape(ncol, nyu)
        t_w, it_w, is_w, mp_w = 1.25, 17, 3, -1.0
        out["wr_meta"] = np.array([t_w, it_w, is_w, mp_w])
        for tagn, grid, arr in (("U", 2, out["spec_u"]), ("V", 1, out["spec_v"]), ("W", 2, out["spec_w"]),
                                ("UV", 2, out["spec_uv"]), ("P", 3, out["spec_p"])):
            fn = os.path.join(td, f"sp{tagn}.dat").encode()
            spc = np.ascontiguousarray(arr, dtype=np.float64)
            ref.ref_write_spectra(ctypes.c_int(grid), _dp(spc), ctypes.c_int(is_w),
                                  ctypes.c_double(t_w), ctypes.c_int(it_w), ctypes.c_double(mp_w),
                                  ctypes.c_int(len(fn)), ctypes.c_char_p(fn))
            out[f"file_sp{tagn}"] = np.frombuffer(open(fn, "rb").read(), np.uint8)
        ref.ref_write_stats(ctypes.c_int(c["ntilex"]), ctypes.c_int(c["ntilez"]), _dp(st), _dp(cb),
                            ctypes.c_int(is_w), ctypes.c_double(t_w), ctypes.c_int(it_w),
                            ctypes.c_double(mp_w), ctypes.c_int(len(d)), ctypes.c_char_p(d))
        fst = [f for f in os.listdir(td) if f.startswith("stats_")][0]
        out["file_stats"] = np.frombuffer(open(os.path.join(td, fst), "rb").read(), np.uint8)
        out["file_stats_name"] = np.frombuffer(fst.encode(), np.uint8)
    # ---- FFTPACK
    rng = np.random.default_rng(7)
    for n in (12, 24, 48, 96):
        x = rng.standard_normal(n)
        ws = np.zeros(2 * n + 18)
        fft.rfti_(ctypes.byref(ctypes.c_int(n)), _dp(ws))
        buf = np.zeros(n + 2)
        buf[:n] = x
        fft.rft_(_dp(buf), ctypes.byref(ctypes.c_int(n + 2)), ctypes.byref(ctypes.c_int(1)),
                 ctypes.byref(ctypes.c_int(-1)), _dp(ws))
        out[f"rft_in_{n}"], out[f"rft_fwd_{n}"] = x, buf.copy()
        back = buf.copy()
        fft.rft_(_dp(back), ctypes.byref(ctypes.c_int(n + 2)), ctypes.byref(ctypes.c_int(1)),
                 ctypes.byref(ctypes.c_int(0)), _dp(ws))
        out[f"rft_inv_{n}"] = back
        z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        wc = np.zeros(10 * n + 19)
        fft.cfti_(ctypes.byref(ctypes.c_int(n)), _dp(wc))
        for sign, tag in ((-1, "fwd"), (0, "inv")):
            zb = np.ascontiguousarray(z, dtype=np.complex128).copy()
            fft.cft_(_dp(zb), ctypes.byref(ctypes.c_int(2)), ctypes.byref(ctypes.c_int(2 * n)),
                     ctypes.byref(ctypes.c_int(1)), ctypes.byref(ctypes.c_int(sign)), _dp(wc))
            out[f"cft_{tag}_{n}"] = zb
        out[f"cft_in_{n}"] = z
    # ---- plane transforms from rft/cft: physical (pz x px) <-> spectral (nz x mx) with 3/2 padding
    nx, nz = c["nx"], c["nz"]
    px, pz, mx = 3 * nx // 2, 3 * nz // 2, nx // 2 + 1
    wsx = np.zeros(2 * px + 18)
    fft.rfti_(ctypes.byref(ctypes.c_int(px)), _dp(wsx))
    wsz = np.zeros(10 * pz + 19)
    fft.cfti_(ctypes.byref(ctypes.c_int(pz)), _dp(wsz))
    ph = rng.standard_normal((pz, px))
    buf = np.zeros((pz, px + 2))
    buf[:, :px] = ph
    fft.rft_(_dp(buf), ctypes.byref(ctypes.c_int(px + 2)), ctypes.byref(ctypes.c_int(pz)),
             ctypes.byref(ctypes.c_int(-1)), _dp(wsx))
    hz = np.ascontiguousarray((buf[:, 0::2] + 1j * buf[:, 1::2])[:, :mx])       # (pz, mx)
    fft.cft_(_dp(hz), ctypes.byref(ctypes.c_int(2 * mx)), ctypes.byref(ctypes.c_int(2)),
             ctypes.byref(ctypes.c_int(mx)), ctypes.byref(ctypes.c_int(-1)), _dp(wsz))
    h = nz // 2
    spec = np.concatenate([hz[:h], hz[pz - h:]], axis=0)                          # (nz, mx)
    out.update(plane_phys=ph, plane_spec=spec)
    # inverse: the spectrum back to physical space (cft inverse along z, rft inverse along x)
    pad = np.zeros((pz, mx), complex)
    pad[:h] = spec[:h]
    pad[pz - h:] = spec[h:]
    pad = np.ascontiguousarray(pad)
    fft.cft_(_dp(pad), ctypes.byref(ctypes.c_int(2 * mx)), ctypes.byref(ctypes.c_int(2)),
             ctypes.byref(ctypes.c_int(mx)), ctypes.byref(ctypes.c_int(0)), _dp(wsz))
    bb = np.zeros((pz, px + 2))
    bb[:, 0:2 * mx:2] = pad.real
    bb[:, 1:2 * mx:2] = pad.imag
    fft.rft_(_dp(bb), ctypes.byref(ctypes.c_int(px + 2)), ctypes.byref(ctypes.c_int(pz)),
             ctypes.byref(ctypes.c_int(0)), _dp(wsx))
    out["plane_back"] = bb[:, :px].copy()
    path = os.path.join(HERE, "ref_small.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    sys.exit(main())
