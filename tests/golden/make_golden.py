"""Write the golden vectors under tests/golden/ from the reference itself.

The reference's hot-path Fortran (f/agauge.f, f/staples.f, f/heat_kpi.f,
f/matmul.f, f/a_force.f, f/ahiggs1-3.f, f/gauss.f, f/renorm.f,
f/u_f_renorm.f, f/landau_gauge.f, f/or_gauge_fixing1.f, f/adresses.f) is
compiled from /root/reference/f in FP64 (-fdefault-real-8) by
oracle/ref/Makefile and driven by oracle/ref/refharness.f.  Each fixture holds
the inputs, the exact uniform stream the reference consumed (captured by
replaying the seeded random_number intrinsic), and the reference's outputs.

    make -C oracle/ref && python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import refrun  # noqa: E402
from lceqcd_amd.lattice import couplings_from_xy  # noqa: E402

# couplings of the shipped runs/inputfile (beta=9, x=0.1, y=0.5) mapped by
# f/paraminit.f:150-153, and the trace-convention production couplings
# (beta_G=9, beta_A=0.35, beta_2=-0.9, beta_4=0.1) in the reference
# normalisation (beta_A/2, beta_2/2, beta_4/4)
REF_INPUTFILE = dict(zip(("beta_a", "beta_2", "beta_4"), couplings_from_xy(9.0, 0.1, 0.5)),
                     beta=9.0)
PRODUCTION = dict(beta=9.0, beta_a=0.175, beta_2=-0.45, beta_4=0.025)
# a moderate-hopping point where every accept/reject branch occurs often
MIXED = dict(beta=9.0, beta_a=0.175, beta_2=0.45, beta_4=0.025)


def rand_conf(shape, seed, asig=0.7):
    rng = np.random.default_rng(seed)
    lx, ly, lz = shape
    n = lx * ly * lz // 2
    u = rng.normal(size=(3, 2, n, 4))
    u /= np.linalg.norm(u, axis=-1, keepdims=True)
    a = rng.normal(size=(2, n, 3)) * asig
    return u, a


def save(name, **arrs):
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **{k: np.asarray(v) for k, v in arrs.items()})
    print("wrote", path, os.path.getsize(path), "bytes")


def coup(c):
    return np.array([c["beta"], c["beta_a"], c["beta_2"], c["beta_4"]])


def main():
    if not refrun.available():
        refrun.build()
    # geometry tables (f/adresses.f) on an anisotropic lattice
    shape = (6, 4, 8)
    u, a = rand_conf(shape, 1)
    r = refrun.run("geom", shape, u, a)
    n = 6 * 4 * 8 // 2
    save("geom_6x4x8", shape=shape, siteup=r[0].reshape(2, 3, n).astype(np.int32) - 1,
         sitedn=r[1].reshape(2, 3, n).astype(np.int32) - 1)

    # staples for every (eo, mu)
    u, a = rand_conf(shape, 2)
    v = np.zeros((3, 2, n, 4))
    for eo in (0, 1):
        for mu in (1, 2, 3):
            r = refrun.run("staples", shape, u, a, arg1=eo, arg2=mu)
            v[mu - 1, eo] = r[0].reshape(3, 2, n, 4)[mu - 1, eo]
    save("staples_6x4x8", shape=shape, u=u, a=a, v=v)

    # a_force for both parities
    f = np.zeros((2, n, 3))
    for eo in (0, 1):
        f[eo] = refrun.run("a_force", shape, u, a, arg1=eo)[0].reshape(n, 3)
    save("a_force_6x4x8", shape=shape, u=u, a=a, force=f)

    # agauge, two calls with measflag=1
    for shp, seed, nrep, cp, tag in (((6, 4, 8), 3, 2, MIXED, "mixed"),
                                     ((4, 4, 4), 4, 2, REF_INPUTFILE, "inputfile"),
                                     ((8, 8, 8), 5, 1, PRODUCTION, "production")):
        u, a = rand_conf(shp, seed)
        r = refrun.run("agauge", shp, u, a, arg1=1, nrep=nrep, **cp)
        assert r[-2][0] == 1.0, "uniform count check failed"
        save("agauge_%dx%dx%d_%s" % (shp + (tag,)), shape=shp, u=u, a=a, couplings=coup(cp),
             uniforms=r[-1], accept=[r[2 * k][0] for k in range(nrep)],
             plaq0=[r[2 * k + 1][0] for k in range(nrep)], u_out=r[2 * nrep])

    # Higgs updates: ahiggs1 (delta 0.4), ahiggs2 / ahiggs3 with imode 0 and 1
    shp = (6, 4, 8)
    for cmd, imode, cp, tag in (("ahiggs1", 0, PRODUCTION, "production"),
                                ("ahiggs1", 0, MIXED, "mixed"),
                                ("ahiggs2", 0, MIXED, "imode0"),
                                ("ahiggs2", 1, MIXED, "imode1"),
                                ("ahiggs3", 0, MIXED, "imode0"),
                                ("ahiggs3", 1, MIXED, "imode1"),
                                ("ahiggs2", 0, REF_INPUTFILE, "inputfile")):
        u, a = rand_conf(shp, 10 + imode, asig=0.4)
        r = refrun.run(cmd, shp, u, a, arg1=1, nrep=2, imode=imode, delta=0.4, **cp)
        assert r[-2][0] == 1.0
        obs = np.array([[r[5 * k + j][0] for j in range(5)] for k in range(2)])
        save("%s_6x4x8_%s" % (cmd, tag), shape=shp, u=u, a=a, couplings=coup(cp), imode=imode,
             delta=0.4, uniforms=r[-1], obs=obs, a_out=r[10])

    # renorm of non-unitary links
    u, a = rand_conf(shp, 20)
    u = u * np.random.default_rng(21).uniform(0.5, 1.5, size=u.shape[:-1] + (1,))
    save("renorm_6x4x8", shape=shp, u=u, u_out=refrun.run("renorm", shp, u, a)[0])

    # overrelaxed gauge fixing: three iterations of the landau_gauge loop body
    u, a = rand_conf(shp, 22)
    r = refrun.run("or_gauge_fixing", shp, u, a, nrep=3, lam=1.7, alpha=1.0)
    save("or_gauge_fixing_6x4x8", shape=shp, u=u, a=a, lam=1.7, alpha=1.0,
         da=[r[k][0] for k in range(3)], u_f=r[3], a_f=r[4])
    # full landau_gauge to da <= 1e-11 (alpha-weighted z as in runs/inputfile)
    for alpha, tag in ((1.0, "iso"), (0.5, "alpha05")):
        r = refrun.run("landau_gauge", (4, 4, 4), *rand_conf((4, 4, 4), 23), lam=1.7,
                       alpha=alpha, limit_or=1e-11)
        u, a = rand_conf((4, 4, 4), 23)
        save("landau_gauge_4x4x4_%s" % tag, shape=(4, 4, 4), u=u, a=a, lam=1.7, alpha=alpha,
             limit_or=1e-11, da1=r[0][0], iter_or1=r[1][0], u_f=r[2], a_f=r[3])

    # reference production sweeps (ahiggs2, ahiggs3, agauge, renorm)
    for shp, seed, nrep in (((4, 4, 4), 30, 2), ((8, 8, 8), 31, 1)):
        u, a = rand_conf(shp, seed, asig=0.4)
        r = refrun.run("sweep", shp, u, a, nrep=nrep, **MIXED)
        assert r[-2][0] == 1.0
        obs = np.array([[r[7 * k + j][0] for j in range(7)] for k in range(nrep)])
        save("sweep_%dx%dx%d" % shp, shape=shp, u=u, a=a, couplings=coup(MIXED), uniforms=r[-1],
             obs=obs, u_out=r[7 * nrep], a_out=r[7 * nrep + 1])

    # standalone heat_kpi and gauss on N sites
    rng = np.random.default_rng(40)
    u, a = rand_conf(shp, 41)
    nn = shp[0] * shp[1] * shp[2] // 2
    b = rng.uniform(0.01, 0.6, nn)
    a0 = rng.uniform(-1, 1, nn)
    r = refrun.run("heat_kpi", shp, u, a, ex1=b, ex2=a0)
    save("heat_kpi_%d" % nn, b=b, a0_old=a0, uniforms=r[-1], v=r[0].reshape(nn, 4),
         nacch=r[1][0])
    R = rng.uniform(0.3, 4.0, nn)
    r = refrun.run("gauss", shp, u, a, ex1=R)
    save("gauss_%d" % nn, r=R, uniforms=r[-1], x=r[0])


if __name__ == "__main__":
    main()
