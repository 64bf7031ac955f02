"""Writes tests/golden/reference_cases.json.gz: inputs and expected outputs of the
reference's own unit and integration tests for the MAC-grid path.

Each case records the test it comes from (file:lines in
junwei-jiang/Hybrid-Fluid-Simulation/TestCase/MPhysics/...), the input fields
that test builds (generated here from the same loops, written out as data),
and the values its EXPECT lines check, with the tolerance it uses
(GRID_SOLVER_EPSILON = 1e-4, MPhysics/Common.h:92).  Expected values are the
reference test's own constants, or the closed forms its EXPECT lines evaluate.

Run: python tests/golden/make_fixtures.py   (no reference files are read)
"""
import gzip
import json
import os

import numpy as np

EPS = 1e-4
FLT_MAX = float(np.finfo(np.float32).max)
HERE = os.path.dirname(os.path.abspath(__file__))
T = "TestCase/MPhysics/"


def cc(res, f):
    nx, ny, nz = res
    a = np.zeros((nz, ny, nx), np.float64)
    for i in range(nz):
        for j in range(ny):
            for k in range(nx):
                a[i, j, k] = f(k, j, i)
    return a


def fc(res, comp, f):
    nx, ny, nz = res
    shape = (nz + (comp == 2), ny + (comp == 1), nx + (comp == 0))
    a = np.zeros(shape, np.float64)
    for i in range(shape[0]):
        for j in range(shape[1]):
            for k in range(shape[2]):
                a[i, j, k] = f(k, j, i)
    return a


def L(a):
    return [float(np.float32(v)) for v in np.asarray(a).ravel()]


# ---------------------------------------------------------------------------
# Pressure solver tests (Euler/UnitTest-PressureSolver/UnitTest-PressureSolver.cu)
def pressure_setup(res, fluid_c, solid_air, fluid_air, vel, svel):
    solid = cc(res, lambda k, j, i: -0.5).ravel()
    for idx in solid_air:
        solid[idx] = 0.5
    fluid = cc(res, lambda k, j, i: j - fluid_c).ravel()
    for idx in fluid_air:
        fluid[idx] = 0.5
    u = fc(res, 0, vel[0]); v = fc(res, 1, vel[1]); w = fc(res, 2, vel[2])
    us = fc(res, 0, svel[0]); vs = fc(res, 1, svel[1]); ws = fc(res, 2, svel[2])
    return dict(solid=L(solid), fluid=L(fluid), u=L(u), v=L(v), w=L(w),
                us=L(us), vs=L(vs), ws=L(ws))


S3 = [13, 16]
S4 = [21, 22, 25, 26, 29, 30, 37, 38, 41, 42, 45, 46]
F4 = [21, 22, 25, 26, 37, 38, 41, 42]
AIR4 = [29, 30, 45, 46]
VEL_KJI = (lambda k, j, i: k, lambda k, j, i: j, lambda k, j, i: i)
SV_100 = (lambda k, j, i: 1, lambda k, j, i: 0, lambda k, j, i: 0)


def pressure_cases():
    out = []
    sp = (10, 20, 30)
    # buildMarkers1 :10-120
    inp = pressure_setup((3, 3, 3), 1.5, S3, [], VEL_KJI, SV_100)
    exp = [2.0] * 27
    exp[13] = 1.0
    exp[16] = 0.0
    out.append(dict(name="PressureSolver_buildMarkers1", source=T + "Euler/UnitTest-PressureSolver/UnitTest-PressureSolver.cu:10-120",
                    res=[3, 3, 3], origin=[0, 0, 0], spacing=sp, dt=1.0, inputs=inp,
                    expect=dict(markers={str(i): v for i, v in enumerate(exp)}), tol=EPS))
    # buildMarkers2 :122-231
    inp = pressure_setup((4, 4, 4), 2.5, S4, [], VEL_KJI, SV_100)
    exp = {}
    for i in range(64):
        exp[str(i)] = 0.0 if i in AIR4 else (1.0 if i in F4 else 2.0)
    out.append(dict(name="PressureSolver_buildMarkers2", source=T + "Euler/UnitTest-PressureSolver/UnitTest-PressureSolver.cu:122-231",
                    res=[4, 4, 4], origin=[0, 0, 0], spacing=sp, dt=1.0, inputs=inp,
                    expect=dict(markers=exp), tol=EPS))
    # buildFdmMatrixA1 :233-333 (A[4c..4c+3])
    inp = pressure_setup((3, 3, 3), 1.5, S3, [], VEL_KJI, SV_100)
    exp = {}
    for i in range(27):
        vals = [0.0025, 0, 0, 0] if i == 13 else [1, 0, 0, 0]
        for q in range(4):
            exp[str(4 * i + q)] = float(vals[q])
    out.append(dict(name="PressureSolver_buildFdmMatrixA1", source=T + "Euler/UnitTest-PressureSolver/UnitTest-PressureSolver.cu:233-333",
                    res=[3, 3, 3], origin=[0, 0, 0], spacing=sp, dt=1.0, inputs=inp,
                    expect=dict(A=exp), tol=EPS))
    # buildFdmMatrixA2 :335-475, dt = 2
    inp = pressure_setup((4, 4, 4), 2.5, S4, [], VEL_KJI, SV_100)
    ix, iy, iz = 2 * 0.01, 2 * 0.0025, 2 * 0.00111111
    A = {21: [ix + iy + iz, -ix, -iy, -iz], 22: [ix + iy + iz, 0, -iy, -iz],
         25: [ix + 2 * iy + iz, -ix, 0, -iz], 26: [ix + 2 * iy + iz, 0, 0, -iz],
         37: [ix + iy + iz, -ix, -iy, 0], 38: [ix + iy + iz, 0, -iy, 0],
         41: [ix + 2 * iy + iz, -ix, 0, 0], 42: [ix + 2 * iy + iz, 0, 0, 0]}
    exp = {}
    for i in range(64):
        if i in A:
            vals = A[i]
        elif i in AIR4:
            continue
        else:
            vals = [1, 0, 0, 0]
        for q in range(4):
            exp[str(4 * i + q)] = float(vals[q])
    out.append(dict(name="PressureSolver_buildFdmMatrixA2", source=T + "Euler/UnitTest-PressureSolver/UnitTest-PressureSolver.cu:335-475",
                    res=[4, 4, 4], origin=[0, 0, 0], spacing=sp, dt=2.0, inputs=inp,
                    expect=dict(A=exp), tol=EPS))
    # buildVectorb1 :477-571, solid u = k + 1
    inp = pressure_setup((3, 3, 3), 1.5, S3, [], VEL_KJI,
                         (lambda k, j, i: k + 1, lambda k, j, i: 0, lambda k, j, i: 0))
    exp = {str(i): (-0.2 if i == 13 else 0.0) for i in range(27)}
    out.append(dict(name="PressureSolver_buildVectorb1", source=T + "Euler/UnitTest-PressureSolver/UnitTest-PressureSolver.cu:477-571",
                    res=[3, 3, 3], origin=[0, 0, 0], spacing=sp, dt=1.0, inputs=inp,
                    expect=dict(b=exp), tol=EPS))
    # buildVectorb2 :573-683
    inp = pressure_setup((4, 4, 4), 2.5, S4, [], VEL_KJI, SV_100)
    bv = {21: -0.26666666, 22: -0.06666666, 25: -0.21666666, 26: -0.01666666,
          37: -0.13333333, 38: 0.06666666, 41: -0.08333333, 42: 0.11666666}
    exp = {str(i): float(bv.get(i, 0.0)) for i in range(64)}
    out.append(dict(name="PressureSolver_buildVectorb2", source=T + "Euler/UnitTest-PressureSolver/UnitTest-PressureSolver.cu:573-683",
                    res=[4, 4, 4], origin=[0, 0, 0], spacing=sp, dt=1.0, inputs=inp,
                    expect=dict(b=exp), tol=EPS))
    # MatrixFreePCG1 :685-793, dt = 2
    xv = {21: -64.10625691, 22: -59.1624367, 25: -40.88984875, 26: -36.39546673,
          37: -40.83756019, 38: -35.8937402, 41: -23.60453116, 42: -19.11014936}
    out.append(dict(name="PressureSolver_MatrixFreePCG1", source=T + "Euler/UnitTest-PressureSolver/UnitTest-PressureSolver.cu:685-793",
                    res=[4, 4, 4], origin=[0, 0, 0], spacing=sp, dt=2.0, inputs=inp,
                    expect=dict(x={str(k): v for k, v in xv.items()}), tol=EPS))
    # PressureSolver1 :795-916, solid velocity (k+1, j+1, i+1), dt = 2
    inp1 = pressure_setup((4, 4, 4), 2.5, S4, [], VEL_KJI,
                          (lambda k, j, i: k + 1, lambda k, j, i: j + 1, lambda k, j, i: i + 1))
    out.append(dict(name="PressureSolver_PressureSolver1", source=T + "Euler/UnitTest-PressureSolver/UnitTest-PressureSolver.cu:795-916",
                    res=[4, 4, 4], origin=[0, 0, 0], spacing=sp, dt=2.0, inputs=inp1,
                    expect=dict(u={"26": 2.0, "28": 4.0, "31": 2.0, "33": 4.0, "46": 2.0, "48": 4.0, "51": 2.0, "53": 4.0},
                                v={"25": 2.0, "26": 2.0, "45": 2.0, "46": 2.0},
                                w={"21": 2.0, "22": 2.0, "25": 2.0, "26": 2.0, "53": 4.0, "54": 4.0, "57": 4.0, "58": 4.0}),
                    tol=EPS))
    # PressureSolver2 :918-1035, dt = 2 (one entry uses 10x the tolerance)
    out.append(dict(name="PressureSolver_PressureSolver2", source=T + "Euler/UnitTest-PressureSolver/UnitTest-PressureSolver.cu:918-1035",
                    res=[4, 4, 4], origin=[0, 0, 0], spacing=sp, dt=2.0, inputs=inp,
                    expect=dict(u={"27": 1.011228, "32": 1.101123, "47": 1.011228, "52": 1.101123},
                                v={"29": -0.3216407, "30": -0.2766933, "33": -1.088985, "34": -0.6395467,
                                   "49": 0.276697, "50": 0.321641, "53": 0.639547, "54": 1.088985},
                                w={"37": 0.4487535, "38": 0.448756, "41": 0.8475453, "42": 0.8476455}),
                    tol=EPS, tol_override={"w:41": 10 * EPS}))
    # PressureSolver3 :1037-1246, spacing 1, dt = 1, u = 4 - k
    inp3 = pressure_setup((4, 4, 4), 2.5, S4, [],
                          (lambda k, j, i: 4 - k, lambda k, j, i: 0, lambda k, j, i: 0), SV_100)
    A = {21: [3, -1, -1, -1], 22: [3, 0, -1, -1], 25: [4, -1, 0, -1], 26: [4, 0, 0, -1],
         37: [3, -1, -1, 0], 38: [3, 0, -1, 0], 41: [4, -1, 0, 0], 42: [4, 0, 0, 0]}
    expA = {}
    for i in range(64):
        if i in A:
            vals = A[i]
        elif i in AIR4:
            continue
        else:
            vals = [1, 0, 0, 0]
        for q in range(4):
            expA[str(4 * i + q)] = float(vals[q])
    bv = {21: -1.0, 22: 1.0, 25: -1.0, 26: 1.0, 37: -1.0, 38: 1.0, 41: -1.0, 42: 1.0}
    xv = {21: -0.45454545, 22: 0.45454545, 25: -0.36363636, 26: 0.36363636,
          37: -0.45454545, 38: 0.45454545, 41: -0.36363636, 42: 0.36363636}
    out.append(dict(name="PressureSolver_PressureSolver3", source=T + "Euler/UnitTest-PressureSolver/UnitTest-PressureSolver.cu:1037-1246",
                    res=[4, 4, 4], origin=[0, 0, 0], spacing=[1, 1, 1], dt=1.0, inputs=inp3,
                    expect=dict(A=expA, b={str(i): float(bv.get(i, 0.0)) for i in range(64)},
                                x={str(i): float(xv.get(i, 0.0)) for i in range(64)},
                                u={"26": 1.0, "28": 1.0, "31": 1.0, "33": 1.0, "46": 1.0, "48": 1.0, "51": 1.0, "53": 1.0,
                                   "27": 1.090909, "32": 1.272727, "47": 1.090909, "52": 1.272727},
                                v={"29": -0.090909, "30": 0.090909, "33": -0.363636, "34": 0.363636,
                                   "49": -0.090909, "50": 0.090909, "53": -0.363636, "54": 0.363636},
                                w={"37": 0.0, "38": 0.0, "41": 0.0, "42": 0.0}),
                    tol=EPS))
    # PressureSolver4 :1248-1467, air below the fluid: nothing moves the fluid
    s4 = S4 + [17, 18, 33, 34]
    inp4 = pressure_setup((4, 4, 4), 2.5, s4, [17, 18, 33, 34],
                          (lambda k, j, i: 0, lambda k, j, i: -9.8, lambda k, j, i: 0),
                          (lambda k, j, i: 0, lambda k, j, i: 0, lambda k, j, i: 0))
    A = {c: [4, 0, 0, 0] for c in F4}
    A[21] = [4, -1, -1, -1]; A[22] = [4, 0, -1, -1]; A[25] = [4, -1, 0, -1]; A[26] = [4, 0, 0, -1]
    A[37] = [4, -1, -1, 0]; A[38] = [4, 0, -1, 0]; A[41] = [4, -1, 0, 0]; A[42] = [4, 0, 0, 0]
    expA = {}
    for i in range(64):
        if i in A:
            vals = A[i]
        elif i in AIR4:
            continue
        else:
            vals = [1, 0, 0, 0]
        for q in range(4):
            expA[str(4 * i + q)] = float(vals[q])
    out.append(dict(name="PressureSolver_PressureSolver4", source=T + "Euler/UnitTest-PressureSolver/UnitTest-PressureSolver.cu:1248-1467",
                    res=[4, 4, 4], origin=[0, 0, 0], spacing=[1, 1, 1], dt=1.0, inputs=inp4,
                    expect=dict(A=expA, b={str(i): 0.0 for i in range(64)},
                                x={str(i): 0.0 for i in range(64)},
                                u={str(i): 0.0 for i in [26, 28, 31, 33, 46, 48, 51, 53, 27, 32, 47, 52]},
                                v={str(i): -9.8 for i in [29, 30, 33, 34, 49, 50, 53, 54]},
                                w={str(i): 0.0 for i in [37, 38, 41, 42]}),
                    tol=EPS))
    return out


# ---------------------------------------------------------------------------
# Semi-Lagrangian tests (Euler/UnitTest-SemiLagrangian/UnitTest-SemiLagrangian.cu)
def semilagrangian_cases():
    out = []
    res = (14, 15, 16)
    sp = (10, 20, 30)
    px = cc(res, lambda k, j, i: k * 10 + 5)
    py = cc(res, lambda k, j, i: j * 20 + 10)
    pz = cc(res, lambda k, j, i: i * 30 + 15)
    K = cc(res, lambda k, j, i: k); J = cc(res, lambda k, j, i: j); I = cc(res, lambda k, j, i: i)

    def bt(name, lines, vel, order, ex, ey, ez):
        out.append(dict(name=name, source=T + "Euler/UnitTest-SemiLagrangian/UnitTest-SemiLagrangian.cu:" + lines,
                        kind="backtrace", res=list(res), origin=[0, 0, 0], spacing=list(sp), dt=5.0,
                        order=order, inputs=dict(u=L(fc(res, 0, vel[0])), v=L(fc(res, 1, vel[1])),
                                                 w=L(fc(res, 2, vel[2])), px=L(px), py=L(py), pz=L(pz)),
                        expect=dict(px=L(ex), py=L(ey), pz=L(ez)), tol=EPS))

    c = lambda val: (lambda k, j, i: val)
    # BackTrace1 :236-323 constant velocity (1,2,3), default RK2
    bt("SemiLagrangian_BackTrace1", "231-321", (c(1), c(2), c(3)), 2, K * 10, J * 20, I * 30)
    # BackTrace2 :325-414 zero velocity
    bt("SemiLagrangian_BackTrace2", "323-413", (c(0), c(0), c(0)), 2, px, py, pz)
    vk = (lambda k, j, i: k, lambda k, j, i: j, lambda k, j, i: i)
    # BackTrace3 :416-505 RK1 in the linear field
    bt("SemiLagrangian_BackTrace3", "415-505", vk, 1, K * 10 + 5 - (K + 0.5) * 5,
       J * 20 + 10 - (J + 0.5) * 5, I * 30 + 15 - (I + 0.5) * 5)
    # BackTrace4 :507-598 RK2
    bt("SemiLagrangian_BackTrace4", "507-598", vk, 2,
       K * 10 + 5 - 5 * (K * 10 + 5 - (K + 0.5) * 5 * 0.5) / 10,
       J * 20 + 10 - 5 * (J * 20 + 10 - (J + 0.5) * 5 * 0.5) / 20,
       I * 30 + 15 - 5 * (I * 30 + 15 - (I + 0.5) * 5 * 0.5) / 30)

    # BackTrace5 :600-703 RK3 (closed form written in that test)
    def rk3(p, h):
        v1 = p / h
        v2 = (p - 0.5 * 5 * v1) / h
        v3 = (p - 0.75 * 5 * v2) / h
        return p - 5 * (2.0 / 9.0 * v1 + 3.0 / 9.0 * v2 + 4.0 / 9.0 * v3)
    bt("SemiLagrangian_BackTrace5", "600-703", vk, 3, rk3(px, 10), rk3(py, 20), rk3(pz, 30))

    # AdvectionSolver1 :705-789: CCS on 14x1x1, field k*100+50, RK2
    r1 = (14, 1, 1)
    f1 = cc(r1, lambda k, j, i: k * 100 + 50)
    K1 = cc(r1, lambda k, j, i: k)
    e1 = 10 * (K1 * 10 + 5 - 5 * (K1 * 10 + 5 - (K1 + 0.5) * 5 * 0.5) / 10)
    e1 = e1.ravel()
    e1[0] = f1.ravel()[0]
    out.append(dict(name="SemiLagrangian_AdvectionSolver1", source=T + "Euler/UnitTest-SemiLagrangian/UnitTest-SemiLagrangian.cu:705-794",
                    kind="advect_cc", res=list(r1), origin=[0, 0, 0], spacing=list(sp), dt=5.0, order=2,
                    inputs=dict(field=L(f1), u=L(fc(r1, 0, vk[0])), v=L(fc(r1, 1, vk[1])), w=L(fc(r1, 2, vk[2]))),
                    expect=dict(field=L(e1)), tol=EPS))
    # AdvectionSolver2 :791-889: CCS on 14x15x16, checked where k,j,i > 0
    f2 = cc(res, lambda k, j, i: k * 10 + 5 + j * 20 + 10 + i * 30 + 15)
    tx = K * 10 + 5 - 5 * (K * 10 + 5 - (K + 0.5) * 5 * 0.5) / 10
    ty = J * 20 + 10 - 5 * (J * 20 + 10 - (J + 0.5) * 5 * 0.5) / 20
    tz = I * 30 + 15 - 5 * (I * 30 + 15 - (I + 0.5) * 5 * 0.5) / 30
    mask = ((K > 0) & (J > 0) & (I > 0)).ravel()
    # the test subtracts its three float terms one by one in float:
    # |((dst - TempX) - TempY) - TempZ| < eps
    parts = [L(t.ravel()) for t in (tx, ty, tz)]
    out.append(dict(name="SemiLagrangian_AdvectionSolver2", source=T + "Euler/UnitTest-SemiLagrangian/UnitTest-SemiLagrangian.cu:796-896",
                    kind="advect_cc", res=list(res), origin=[0, 0, 0], spacing=list(sp), dt=5.0, order=2,
                    inputs=dict(field=L(f2), u=L(fc(res, 0, vk[0])), v=L(fc(res, 1, vk[1])), w=L(fc(res, 2, vk[2]))),
                    expect=dict(field_parts={str(i): [parts[0][i], parts[1][i], parts[2][i]] for i in np.nonzero(mask)[0]}),
                    tol=EPS))
    # AdvectionSolver3 :891-1001: CCV -> each component is a CCS advection
    for comp, (src, ex) in enumerate(((px, tx), (py, ty), (pz, tz))):
        out.append(dict(name="SemiLagrangian_AdvectionSolver3_c%d" % comp,
                        source=T + "Euler/UnitTest-SemiLagrangian/UnitTest-SemiLagrangian.cu:898-1007",
                        kind="advect_cc", res=list(res), origin=[0, 0, 0], spacing=list(sp), dt=5.0, order=2,
                        inputs=dict(field=L(src), u=L(fc(res, 0, vk[0])), v=L(fc(res, 1, vk[1])), w=L(fc(res, 2, vk[2]))),
                        expect=dict(field={str(i): float(ex.ravel()[i]) for i in np.nonzero(mask)[0]}), tol=EPS))
    # AdvectionSolver4 :1003-1111: FCV advection, RK2
    iu = fc(res, 0, lambda k, j, i: k * 10); iv = fc(res, 1, lambda k, j, i: j * 20)
    iw = fc(res, 2, lambda k, j, i: i * 30)
    eu = fc(res, 0, lambda k, j, i: k * 10 - 5 * (k * 10 - k * 5 * 0.5) / 10)
    ev = fc(res, 1, lambda k, j, i: j * 20 - 5 * (j * 20 - j * 5 * 0.5) / 20)
    ew = fc(res, 2, lambda k, j, i: i * 30 - 5 * (i * 30 - i * 5 * 0.5) / 30)
    out.append(dict(name="SemiLagrangian_AdvectionSolver4", source=T + "Euler/UnitTest-SemiLagrangian/UnitTest-SemiLagrangian.cu:1009-1111",
                    kind="advect_fc", res=list(res), origin=[0, 0, 0], spacing=list(sp), dt=5.0, order=2,
                    inputs=dict(iu=L(iu), iv=L(iv), iw=L(iw), u=L(fc(res, 0, vk[0])), v=L(fc(res, 1, vk[1])), w=L(fc(res, 2, vk[2]))),
                    expect=dict(u=L(eu), v=L(ev), w=L(ew)), tol=EPS))
    return out


# ---------------------------------------------------------------------------
def external_forces_cases():
    res = (14, 15, 16)
    u = fc(res, 0, lambda k, j, i: k); v = fc(res, 1, lambda k, j, i: j); w = fc(res, 2, lambda k, j, i: i)
    steps = []
    for n in range(1, 10):
        steps.append(dict(u=L(u + 10 * n * 2), v=L(v - 9.8 * n * 2), w=L(w + 50 * n * 2)))
    return [dict(name="ExternalForcesSolver_ApplyExternalForces",
                 source=T + "Euler/UnitTest-ExternalForces/UnitTest-ExternalForces.cu:10-149",
                 res=list(res), dt=2.0, gravity=[0, -9.8, 0], force=[10, 0, 50],
                 inputs=dict(u=L(u), v=L(v), w=L(w)), expect_steps=steps,
                 expect_after_reset=[dict(u=L(u), v=L(v - 9.8 * n * 2), w=L(w)) for n in range(1, 10)],
                 tol=EPS)]


# ---------------------------------------------------------------------------
# Extrapolation tests (Euler/UnitTest-Extrapolation/UnitTest-Extrapolation.cu)
def extrapolation_cases():
    out = []
    U = FLT_MAX
    src = T + "Euler/UnitTest-Extrapolation/UnitTest-Extrapolation.cu:"

    def arr(n, sets):
        a = [U] * n
        for k, val in sets.items():
            a[k] = val
        return a

    # buildDisMarkers1 + extrapolationData1 :59-125 / :402-445 (6x6x1, 3 layers)
    m1 = [U, U, 2, 2, U, U, U, 2, 1, 1, 2, U, 2, 1, 0, 0, 1, 2, 2, 1, 0, 1, 2, U,
          U, 2, 1, 2, U, U, U, U, 2, U, U, U]
    d1 = [U, U, 0, 1, U, U, U, 0, 0, 1, 1, U, 0, 0, 0, 1, 1, 1, 2, 2, 2, 1.5, 1.25, U,
          U, 2, 2, 1.75, U, U, U, U, 2, U, U, U]
    out.append(dict(name="Extrapolation_1", source=src + "71-129,383-441", kind="grid", shape=[1, 6, 6], layers=3,
                    inputs=dict(data=arr(36, {14: 0, 15: 1, 20: 2})), expect=dict(layers=m1, data=d1), tol=EPS))
    # buildDisMarkers2 + extrapolationData2 :127-170 / :447-507 (4x4x4, 2 layers)
    ones = [1, 4, 6, 9, 10, 17, 18, 20, 23, 24, 27, 29, 30, 37, 38, 41, 42]
    zeros = [5, 21, 22, 25, 26]
    m2 = [0 if i in zeros else (1 if i in ones else U) for i in range(64)]
    dv = {1: 0, 4: 0, 6: 0.5, 9: -1, 10: 2, 17: 0, 18: 1, 20: 0, 23: 1, 24: -2, 27: 2, 29: -2,
          30: 2, 37: 0, 38: 1, 41: -2, 42: 2, 5: 0, 21: 0, 22: 1, 25: -2, 26: 2}
    d2 = [float(dv[i]) if i in dv else U for i in range(64)]
    out.append(dict(name="Extrapolation_2", source=src + "131-170,443-506", kind="grid", shape=[4, 4, 4], layers=2,
                    inputs=dict(data=arr(64, {5: 0, 21: 0, 22: 1, 25: -2, 26: 2})),
                    expect=dict(layers=m2, data=d2), tol=EPS))
    # buildDisMarkers3 + extrapolationData3 :172-199 / :509-537 (3x3x1, 200 layers)
    out.append(dict(name="Extrapolation_3", source=src + "172-202,508-538", kind="grid", shape=[1, 3, 3], layers=200,
                    inputs=dict(data=arr(9, {2: 0, 6: 1})),
                    expect=dict(layers=[2, 1, 0, 1, 2, 1, 0, 1, 2], data=[0.5, 0, 0, 1, 0.5, 0, 1, 1, 0.5]), tol=EPS))
    # buildDisMarkers4 + extrapolationData4 :201-285 / :539-631 (FCV on 4x4x4, 3 layers)
    mx = {0: 0, 1: 0, 2: 1, 3: 2, 5: 1, 6: 1, 7: 2, 10: 2, 11: 2, 20: 1, 21: 1, 22: 2, 25: 2, 26: 2, 40: 2, 41: 2}
    dx = {0: 1, 1: 2, 2: 2, 3: 2, 5: 1, 6: 2, 7: 2, 10: 1, 11: 2, 20: 1, 21: 2, 22: 2, 25: 1, 26: 2, 40: 1, 41: 2}
    my = {0: 0, 1: 1, 2: 2, 4: 1, 5: 2, 8: 2, 20: 1, 21: 2, 24: 2, 40: 2}
    mz = {0: 0, 1: 1, 2: 2, 4: 1, 5: 2, 8: 2, 16: 1, 17: 2, 20: 2, 32: 2}
    nxf, nyf, nzf = 5 * 4 * 4, 4 * 5 * 4, 4 * 4 * 5
    out.append(dict(name="Extrapolation_4_FCV", source=src + "204-293,540-629", kind="fc", res=[4, 4, 4], layers=3,
                    inputs=dict(u=arr(nxf, {0: 1, 1: 2}), v=arr(nyf, {0: 3}), w=arr(nzf, {0: 4})),
                    expect=dict(layers_u=[mx.get(i, U) for i in range(nxf)], layers_v=[my.get(i, U) for i in range(nyf)],
                                layers_w=[mz.get(i, U) for i in range(nzf)],
                                u=[float(dx[i]) if i in dx else U for i in range(nxf)],
                                v=[3.0 if i in my else U for i in range(nyf)],
                                w=[4.0 if i in mz else U for i in range(nzf)]), tol=EPS))
    # buildDisMarkers5 + extrapolationData5 :287-400 / :633-716 (CCV = three CC grids)
    mx5 = {0: 0, 1: 0, 2: 1, 3: 2, 4: 1, 5: 1, 6: 2, 8: 2, 9: 2, 16: 1, 17: 1, 18: 2, 20: 2, 21: 2, 32: 2, 33: 2}
    dx5 = {0: 1, 1: 2, 2: 2, 3: 2, 4: 1, 5: 2, 6: 2, 8: 1, 9: 2, 16: 1, 17: 2, 18: 2, 20: 1, 21: 2, 32: 1, 33: 2}
    myz = {0: 0, 1: 1, 2: 2, 4: 1, 5: 2, 8: 2, 16: 1, 17: 2, 20: 2, 32: 2}
    out.append(dict(name="Extrapolation_5_CCV_x", source=src + "295-381,631-716", kind="grid", shape=[4, 4, 4], layers=3,
                    inputs=dict(data=arr(64, {0: 1, 1: 2})),
                    expect=dict(layers=[mx5.get(i, U) for i in range(64)],
                                data=[float(dx5[i]) if i in dx5 else U for i in range(64)]), tol=EPS))
    for nm, val in (("y", 3.0), ("z", 4.0)):
        out.append(dict(name="Extrapolation_5_CCV_" + nm, source=src + "295-381,631-716", kind="grid", shape=[4, 4, 4], layers=3,
                        inputs=dict(data=arr(64, {0: val})),
                        expect=dict(layers=[myz.get(i, U) for i in range(64)],
                                    data=[val if i in myz else U for i in range(64)]), tol=EPS))
    return out


# ---------------------------------------------------------------------------
def boundary_cases():
    out = []
    src = T + "Euler/UnitTest-BoundarySDF/UnitTest-BoundarySDF.cu:"
    origin = [1.0, 3.0, -10.0]
    spacing = [11.0, 3.40, -140.0]

    def sdfs(res):
        nx, ny, nz = res
        a = np.full((nz, ny, nx), 100.0); b = np.full((nz, ny, nx), 200.0)
        for i in range(nz):
            for j in range(ny):
                for k in range(nx):
                    if k in (0, nx - 1) or j in (0, ny - 1) or i in (0, nz - 1):
                        a[i, j, k] = -1; b[i, j, k] = -2
                    elif k in (1, nx - 2) or j in (1, ny - 2) or i in (1, nz - 2):
                        b[i, j, k] = -2
        return a, b

    # updateTotalBoundarysSDFAndMarker :126-203
    res = (4, 4, 4)
    a, b = sdfs(res)
    exp = {}
    for i in range(4):
        for j in range(4):
            for k in range(4):
                idx = (i * 4 + j) * 4 + k
                if k in (0, 3) or j in (0, 3) or i in (0, 3):
                    exp[str(idx)] = -1.0
                elif k in (1, 2) or j in (1, 2) or i in (1, 2):
                    exp[str(idx)] = -2.0
    out.append(dict(name="GenerateBoundarySDF_updateTotalBoundarysSDFAndMarker", source=src + "131-203",
                    res=list(res), origin=origin, spacing=spacing,
                    solids=[dict(sdf=L(a), vel=[0, 0, 0]), dict(sdf=L(b), vel=[0, 0, 0])],
                    expect=dict(marker=exp), tol=EPS))
    # updateTotalBoundarysVel :205-317 (faces shared by the two solids are
    # checked up to sign, |v| = 2,3,4, because the reference writes them racily)
    res = (12, 44, 14)
    a, b = sdfs(res)
    nx, ny, nz = res
    ex = {"abs": {}, "val": {}}
    for comp, (nm, sx, sy, sz, mag) in enumerate((("u", 1, 0, 0, 2.0), ("v", 0, 1, 0, 3.0), ("w", 0, 0, 1, 4.0))):
        X, Y, Z = nx + sx, ny + sy, nz + sz
        for i in range(Z):
            for j in range(Y):
                for k in range(X):
                    idx = (i * Y + j) * X + k
                    ring0 = [k in (0, X - 1), j in (0, Y - 1), i in (0, Z - 1)]
                    ring1 = [k in (1, X - 2), j in (1, Y - 2), i in (1, Z - 2)]
                    # the component's own axis uses two outer layers (faces of the ring-0 cells)
                    ax = comp
                    pos = (k, j, i)[ax]
                    n_ax = (X, Y, Z)[ax]
                    ring0[ax] = pos in (0, 1, n_ax - 1, n_ax - 2)
                    ring1[ax] = pos in (2, n_ax - 3)
                    if any(ring0):
                        ex["abs"][nm + ":" + str(idx)] = mag
                    elif any(ring1):
                        ex["val"][nm + ":" + str(idx)] = -mag
    # moveBoundarys :319-435 (translation after one update = -dt * velocity)
    for dt in (0.0, 1.0):
        out.append(dict(name="GenerateBoundarySDF_moveBoundarys_dt%g" % dt, source=src + "319-435",
                        kind="move", res=list(res), origin=[0, 0, 0], spacing=[1, 1, 1], dt=dt,
                        solids=[dict(sdf=L(a), vel=[2, 3, 4]), dict(sdf=L(b), vel=[-2, -3, -4])],
                        expect=dict(translation=[[-dt * 2, -dt * 3, -dt * 4], [dt * 2, dt * 3, dt * 4]]), tol=EPS))
    # reSamplingCurrentBoundarysSDF :437-470 (dt = 0: the re-sample returns the SDF)
    r4 = (4, 4, 4)
    a4, b4 = sdfs(r4)
    out.append(dict(name="GenerateBoundarySDF_reSamplingCurrentBoundarysSDF", source=src + "437-470",
                    kind="resample", res=list(r4), origin=origin, spacing=spacing, dt=0.0,
                    solids=[dict(sdf=L(a4), vel=[2, 3, 4]), dict(sdf=L(b4), vel=[-2, -3, -4])],
                    expect=dict(sdf=[L(a4), L(b4)]), tol=EPS))
    out.append(dict(name="GenerateBoundarySDF_updateTotalBoundarysVel", source=src + "205-317",
                    res=list(res), origin=origin, spacing=spacing,
                    solids=[dict(sdf=L(a), vel=[2, 3, 4]), dict(sdf=L(b), vel=[-2, -3, -4])],
                    expect=ex, tol=EPS))
    return out


# ---------------------------------------------------------------------------
def cg_cases():
    src = T + "Math/UnitTest-CuMatrixFreePCG/UnitTest-CuMatrixFreePCG.cpp:"
    A3 = [[4, -2, -1], [-2, 4, -2], [-1, -2, 3]]
    rows = [0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 6, 6, 6, 6, 7, 7, 7, 7]
    cols = [0, 1, 2, 4, 0, 1, 3, 5, 0, 2, 3, 6, 1, 2, 3, 7, 0, 4, 5, 6, 1, 4, 5, 7, 2, 4, 6, 7, 3, 5, 6, 7]
    vals = [3, -1, -1, -1, -1, 3, -1, -1, -1, 3, -1, -1, -1, -1, 3, -1, -1, 3, -1, -1, -1, -1, 3, -1, -1, -1, 3, -1, -1, -1, -1, 1]
    A8 = np.zeros((8, 8))
    for r, c, v in zip(rows, cols, vals):
        A8[r, c] = v
    return [dict(name="CuMatrixFreePCG_Main", source=src + "29-74", A=A3, minv=[0.25, 0.25, 1 / 3.0],
                 b=[0, -2, 3], max_it=1000, repeats=1, residual_max=1e-6),
            dict(name="CuMatrixFreePCG_Main2", source=src + "76-125", A=A8.tolist(),
                 minv=[1 / 3.0] * 7 + [1.0], b=[-8, -16, -55, -28, -110, 52, 55, -33], max_it=8,
                 repeats=500, residual_max=1e-4)]


# ---------------------------------------------------------------------------
def field_cases():
    out = []
    src = T + "DataStructure/UnitTest-FaceCenteredVectorGrid/UnitTest-FaceCenteredVectorField.cu:"
    res = (2, 2, 1)
    su = fc(res, 0, lambda k, j, i: k % 2); sv = fc(res, 1, lambda k, j, i: j % 2); sw = fc(res, 2, lambda k, j, i: i % 2)
    pos = {}
    pos["u"] = (fc(res, 0, lambda k, j, i: k), fc(res, 0, lambda k, j, i: j + 0.5), fc(res, 0, lambda k, j, i: i + 0.5))
    pos["v"] = (fc(res, 1, lambda k, j, i: k + 0.5), fc(res, 1, lambda k, j, i: j), fc(res, 1, lambda k, j, i: i + 0.5))
    pos["w"] = (fc(res, 2, lambda k, j, i: k + 0.5), fc(res, 2, lambda k, j, i: j + 0.5), fc(res, 2, lambda k, j, i: i))
    out.append(dict(name="FaceCenteredVectorField_CuSamplingTrilinear", source=src + "436-537",
                    kind="sample_fc", res=list(res), origin=[0, 0, 0], spacing=[1, 1, 1],
                    inputs=dict(u=L(su), v=L(sv), w=L(sw), **{"p" + c + a: L(pos[c][q]) for c in "uvw" for q, a in enumerate("xyz")}),
                    expect=dict(u=[0, 1, 0, 0, 1, 0], v=[0, 0, 1, 1, 0, 0], w=[0, 0, 0, 0, 1, 1, 1, 1]), tol=EPS))
    # FieldDivergence :2304-2500 (three sub-cases)
    r = (2, 10, 20)
    sp = [10, 20, 30]
    og = [-10, -40, -50]
    u = fc(r, 0, lambda k, j, i: k % 2); v = fc(r, 1, lambda k, j, i: j % 3); w = fc(r, 2, lambda k, j, i: i % 4)
    e = cc(r, lambda k, j, i: ((k + 1) % 2 - k % 2) / 10.0 + ((j + 1) % 3 - j % 3) / 20.0 + ((i + 1) % 4 - i % 4) / 30.0)
    out.append(dict(name="FaceCenteredVectorField_FieldDivergence_a", source=src + "2304-2366", kind="divergence",
                    res=list(r), origin=og, spacing=sp, inputs=dict(u=L(u), v=L(v), w=L(w)), expect=dict(div=L(e)), tol=EPS))
    u = fc(r, 0, lambda k, j, i: k); v = fc(r, 1, lambda k, j, i: j); w = fc(r, 2, lambda k, j, i: i)
    out.append(dict(name="FaceCenteredVectorField_FieldDivergence_b", source=src + "2368-2427", kind="divergence",
                    res=list(r), origin=og, spacing=sp, inputs=dict(u=L(u), v=L(v), w=L(w)),
                    expect=dict(div=[0.1 + 0.05 + 0.03333333] * 400), tol=EPS))
    r = (2, 2, 2)
    uu = [15, 20, 5, 8, 7, 22, 11, 94, 45, 33, 0, 12]
    vv = [5, 0, 25, 18, 67, 42, 38, 49, 45, 42, 41, 42]
    ww = [51, 2, 4, 11, 34, 15, 18, 0, 54, 19, 0, 21]
    e = [5 / 10 + 20 / 20 - 17 / 30, -15 / 10 + 18 / 20 + 13 / 30, -1 / 10 + 42 / 20 + 14 / 30,
         15 / 10 + 24 / 20 - 11 / 30, 83 / 10 + 7 / 20 + 20 / 30, -49 / 10 - 7 / 20 + 4 / 30,
         -33 / 10 - 4 / 20 - 18 / 30, 12 / 10 + 0 / 20 + 21 / 30]
    out.append(dict(name="FaceCenteredVectorField_FieldDivergence_c", source=src + "2429-2497", kind="divergence",
                    res=list(r), origin=og, spacing=sp, inputs=dict(u=uu, v=vv, w=ww), expect=dict(div=e), tol=EPS))
    return out


# ---------------------------------------------------------------------------
def integration_cases():
    src = T + "Euler/IntegrationTest-GridFluidSolver/IntegrationTest-GridFluidSolver.cpp:"
    res = (4, 4, 4)
    solid = cc(res, lambda k, j, i: -0.5).ravel()
    for idx in S4:
        solid[idx] = 0.5
    fluid = cc(res, lambda k, j, i: j - 2.5).ravel()
    c1 = dict(name="GridFluidSolver1", source=src + "228-311", res=list(res), origin=[0, 0, 0], spacing=[1, 1, 1],
              dt=1.0, steps=3, inputs=dict(solid=L(solid), fluid=L(fluid)),
              expect=dict(all_faces_zero=True), tol=EPS)
    solid2 = solid.copy()
    fluid2 = fluid.copy()
    for idx in (17, 18, 33, 34):
        solid2[idx] = 0.5
        fluid2[idx] = 0.5
    faces = [25, 26, 29, 30, 33, 34, 45, 46, 49, 50, 53, 54]
    c2 = dict(name="GridFluidSolver2", source=src + "313-401", res=list(res), origin=[0, 0, 0], spacing=[1, 1, 1],
              dt=0.01, steps=10, inputs=dict(solid=L(solid2), fluid=L(fluid2)),
              expect=dict(v_faces=faces, v_per_step=[-9.8 * 0.01 * (n + 1) for n in range(10)]), tol=EPS)
    return [c1, c2]


# ---------------------------------------------------------------------------
# Particle transfers (Euler/UnitTest-ParticlesAdvection/UnitTest-ParticlesAdvection.cu).
PT = T + "Euler/UnitTest-ParticlesAdvection/UnitTest-ParticlesAdvection.cu"
A_RES, A_O, A_H = (4, 4, 1), (0.0, 20.0, -30.0), (30.0, 20.0, 10.0)
B_RES = (1, 4, 4)
FIELD_4x4 = [64.0, 64.0, 64.0, 0.0, 64.0, 64.0, 96.0, 128.0, 64.0, 96.0, 128.0, 128.0, 0.0, 128.0, 128.0, 128.0]
FIELD_4x4_B = [36.0, 36.0, 36.0, 0.0, 36.0, 36.0, 54.0, 72.0, 36.0, 54.0, 72.0, 72.0, 0.0, 72.0, 72.0, 72.0]


def cells_particles(res, origin, spacing, cells, per_cell):
    """Particles inside the given cells, per_cell each, in the order of
    CEulerParticles::generateParticlesInFluid (cells x-fastest, then the
    particles of a cell).  The tests that use generated particles check values
    that do not depend on where in its cell a particle lies; the fractions
    here are fixed instead of the reference's rand()."""
    nx, ny, _ = res
    pos = []
    for c in sorted(cells):
        x, y, z = c % nx, (c // nx) % ny, c // (nx * ny)
        for i in range(per_cell):
            f = [(0.13 + 0.71 * ((i * 3 + a) % 7) / 7.0) for a in range(3)]
            pos += [(x + f[0]) * spacing[0] + origin[0], (y + f[1]) * spacing[1] + origin[1],
                    (z + f[2]) * spacing[2] + origin[2]]
    return pos


def particle_cases():
    out = []
    eight = [21, 22, 25, 26, 37, 38, 41, 42]
    o8, h8 = (-10.0, -20.0, -30.0), (10.0, 20.0, 30.0)
    # P2G to a cell-centred scalar field
    pos = cells_particles((4, 4, 4), o8, h8, eight, 8)
    out.append(dict(name="p2g_ccs_sum1", src=PT + ":116-196", kind="p2g", field=0, alg="p2gsum", res=[4, 4, 4],
                    origin=list(o8), spacing=list(h8), pos=pos, val=[10.0] * 64, span=1,
                    expect={str(i): (80.0 if i in eight else 0.0) for i in range(64)}, tol=EPS))
    pos = cells_particles((2, 2, 1), A_O, A_H, [0], 5)
    out.append(dict(name="p2g_ccs_sum2", src=PT + ":198-251", kind="p2g", field=0, alg="p2gsum", res=[2, 2, 1],
                    origin=list(A_O), spacing=list(A_H), pos=pos, val=[1.0, 2.0, 3.0, 4.0, 5.0], span=1,
                    expect={"0": 15.0, "1": 0.0, "2": 0.0, "3": 0.0}, tol=EPS))
    posA = [45.0, 50.0, -25.0, 75.0, 70.0, -15.0]
    posB = [15.0, 60.0, -10.0, 15.0, 60.0, -10.0]
    e = {str(i): 0.0 for i in range(16)}
    e.update({"5": 64.0, "10": 128.0})
    out.append(dict(name="p2g_ccs_linear1", src=PT + ":253-305", kind="p2g", field=0, alg="linear",
                    res=list(A_RES), origin=list(A_O), spacing=list(A_H), pos=posA, val=[64.0, 128.0], span=1,
                    expect=e, tol=EPS))
    e = {str(i): (96.0 if i in (5, 6, 9, 10) else 0.0) for i in range(16)}
    out.append(dict(name="p2g_ccs_linear2", src=PT + ":307-354", kind="p2g", field=0, alg="linear",
                    res=list(B_RES), origin=list(A_O), spacing=list(A_H), pos=posB, val=[64.0, 128.0], span=1,
                    expect=e, tol=EPS))
    e = {"0": 64.0, "1": 64.0, "2": 64.0, "4": 64.0, "5": 38.0 * 64.0 / 37.0, "6": 96.0, "7": 128.0,
         "8": 64.0, "9": 96.0, "10": 73.0 * 64.0 / 37.0, "11": 128.0, "13": 128.0, "14": 128.0, "15": 128.0}
    out.append(dict(name="p2g_ccs_quadratic1", src=PT + ":356-416", kind="p2g", field=0, alg="quadratic",
                    res=list(A_RES), origin=list(A_O), spacing=list(A_H), pos=posA, val=[64.0, 128.0], span=1,
                    expect=e, tol=EPS))
    e = {str(i): (96.0 if i in (5, 6, 9, 10) else 0.0) for i in range(16)}
    out.append(dict(name="p2g_ccs_quadratic2", src=PT + ":418-465", kind="p2g", field=0, alg="quadratic",
                    res=list(B_RES), origin=list(A_O), spacing=list(A_H), pos=posB, val=[64.0, 128.0], span=1,
                    expect=e, tol=EPS))
    e = {"0": 64.0, "1": 64.0, "2": 64.0, "4": 64.0, "5": (64.0 * 16.0 + 128.0) / 17.0, "6": 96.0, "7": 128.0,
         "8": 64.0, "9": 96.0, "10": (64.0 + 128.0 * 16.0) / 17.0, "11": 128.0, "13": 128.0, "14": 128.0,
         "15": 128.0}
    out.append(dict(name="p2g_ccs_cubic1", src=PT + ":467-519", kind="p2g", field=0, alg="cubic",
                    res=list(A_RES), origin=list(A_O), spacing=list(A_H), pos=posA, val=[64.0, 128.0], span=1,
                    expect=e, tol=EPS))
    out.append(dict(name="p2g_ccs_cubic2", src=PT + ":521-562", kind="p2g", field=0, alg="cubic",
                    res=list(B_RES), origin=list(A_O), spacing=list(A_H), pos=posB, val=[64.0, 128.0], span=1,
                    expect={str(i): 96.0 for i in range(16)}, tol=EPS))
    # G2P from a cell-centred scalar field
    pos = cells_particles((4, 4, 4), o8, h8, eight, 8)
    out.append(dict(name="g2p_ccs_nearest1", src=PT + ":1771-1840", kind="g2p", field=0, alg="g2pnearest",
                    res=[4, 4, 4], origin=list(o8), spacing=list(h8), pos=pos, grid=[float(i) for i in range(64)],
                    span=1, expect={str(8 * k + q): float(c) for k, c in enumerate(eight) for q in range(8)},
                    tol=EPS))
    for name, src, res, ps, fld, alg, ev in [
            ("g2p_ccs_linear1", "1842-1894", A_RES, [45.0, 50.0, -24.0, 75.0, 70.0, -26.0], FIELD_4x4, "linear",
             [64.0, 128.0]),
            ("g2p_ccs_linear2", "1896-1948", B_RES, [15.0, 60.0, -10.0, 15.0, 70.0, -5.0], FIELD_4x4, "linear",
             [96.0, 128.0]),
            ("g2p_ccs_quadratic1", "1950-2010", A_RES, [45.0, 50.0, -24.0, 75.0, 70.0, -26.0], FIELD_4x4,
             "quadratic", [71.0, 121.0]),
            ("g2p_ccs_quadratic2", "2012-2065", B_RES, [15.0, 50.0, -15.0, 15.0, 70.0, -5.0], FIELD_4x4,
             "quadratic", [71.0, 121.0]),
            ("g2p_ccs_cubic1", "2067-2120", A_RES, [45.0, 50.0, -24.0, 75.0, 70.0, -26.0], FIELD_4x4_B, "cubic",
             [41.0, 67.0]),
            ("g2p_ccs_cubic2", "2122-2175", B_RES, [15.0, 50.0, -15.0, 15.0, 70.0, -5.0], FIELD_4x4_B, "cubic",
             [41.0, 67.0])]:
        out.append(dict(name=name, src=PT + ":" + src, kind="g2p", field=0, alg=alg, res=list(res),
                        origin=list(A_O), spacing=list(A_H), pos=ps, grid=fld, span=1,
                        expect={"0": ev[0], "1": ev[1]}, tol=EPS))
    # P2G to a face-centred vector field (each particle's 3 components to the 3 face grids)
    o8, h8 = (-10.0, -20.0, -30.0), (10.0, 20.0, 30.0)
    base = (o8[0] + 1.25 * h8[0], o8[1] + 1.25 * h8[1], -16.0)
    pos = []
    for y in range(4):
        for x in range(4):
            pos += [base[0] + x * 0.5 * h8[0], base[1] + y * 0.5 * h8[1], base[2]]
    ex = {str(i): (2.0 if i in (6, 8, 11, 13) else 4.0 if i in (7, 12) else 0.0) for i in range(20)}
    ey = {str(i): (4.0 if i in (5, 6, 13, 14) else 8.0 if i in (9, 10) else 0.0) for i in range(20)}
    ez = {str(i): (12.0 if i in (5, 6, 9, 10) else 0.0) for i in range(32)}
    out.append(dict(name="p2g_fcv_sum1", src=PT + ":1074-1189", kind="p2g_fcv", alg="p2gsum", res=[4, 4, 1],
                    origin=list(o8), spacing=list(h8), pos=pos, val=[1.0, 2.0, 3.0] * 16,
                    expect=dict(x=ex, y=ey, z=ez), tol=EPS))
    base = (-6.0, o8[1] + 1.25 * h8[1], o8[2] + 1.25 * h8[2])
    pos = []
    for z in range(4):
        for y in range(4):
            pos += [base[0], base[1] + y * 0.5 * h8[1], base[2] + z * 0.5 * h8[2]]
    ex = {str(i): (4.0 if i in (10, 12, 18, 20) else 0.0) for i in range(32)}
    ey = {str(i): (4.0 if i in (6, 8, 11, 13) else 8.0 if i in (7, 12) else 0.0) for i in range(20)}
    ez = {str(i): (6.0 if i in (5, 6, 13, 14) else 12.0 if i in (9, 10) else 0.0) for i in range(20)}
    out.append(dict(name="p2g_fcv_sum2", src=PT + ":1191-1306", kind="p2g_fcv", alg="p2gsum", res=[1, 4, 4],
                    origin=list(o8), spacing=list(h8), pos=pos, val=[1.0, 2.0, 3.0] * 16,
                    expect=dict(x=ex, y=ey, z=ez), tol=EPS))
    vel = [1.0, 3.0, 5.0, 2.0, 4.0, 6.0]
    p1 = [30.0, 50.0, -24.0, 60.0, 70.0, -26.0]
    p2 = [45.0, 40.0, -24.0, 75.0, 60.0, -26.0]
    p3 = [15.0, 50.0, -20.0, 15.0, 70.0, -10.0]

    def sparse(n, d):
        e = {str(i): 0.0 for i in range(n)}
        e.update({str(k): v for k, v in d.items()})
        return e
    out.append(dict(name="p2g_fcv_linear1", src=PT + ":1308-1354", kind="p2g_fcv", alg="linear",
                    res=list(A_RES), origin=list(A_O), spacing=list(A_H), pos=p1, val=vel,
                    expect=dict(x=sparse(20, {6: 1.0, 12: 2.0})), tol=EPS))
    out.append(dict(name="p2g_fcv_linear2", src=PT + ":1356-1400", kind="p2g_fcv", alg="linear",
                    res=list(A_RES), origin=list(A_O), spacing=list(A_H), pos=p2, val=vel,
                    expect=dict(y=sparse(20, {5: 3.0, 10: 4.0})), tol=EPS))
    out.append(dict(name="p2g_fcv_linear3", src=PT + ":1402-1446", kind="p2g_fcv", alg="linear",
                    res=list(B_RES), origin=list(A_O), spacing=list(A_H), pos=p3, val=vel,
                    expect=dict(z=sparse(20, {5: 5.0, 10: 6.0})), tol=EPS))

    def quad(a, b):
        return {"0": a, "1": a, "2": a, "4": a, "5": (9.0 / 16.0 * a + 1.0 / 64.0 * b) * 64.0 / 37.0,
                "6": (a + b) / 2, "7": b, "8": a, "9": (a + b) / 2,
                "10": (1.0 / 64.0 * a + 9.0 / 16.0 * b) * 64.0 / 37.0, "11": b, "13": b, "14": b, "15": b}
    qx = {"0": 1.0, "1": 1.0, "2": 1.0, "5": 1.0, "6": (9.0 / 16.0 * 1.0 + 1.0 / 64.0 * 2.0) * 64.0 / 37.0,
          "7": 1.5, "8": 2.0, "10": 1.0, "11": 1.5, "12": (1.0 / 64.0 * 1.0 + 9.0 / 16.0 * 2.0) * 64.0 / 37.0,
          "13": 2.0, "16": 2.0, "17": 2.0, "18": 2.0}
    out.append(dict(name="p2g_fcv_quadratic1", src=PT + ":1448-1500", kind="p2g_fcv", alg="quadratic",
                    res=list(A_RES), origin=list(A_O), spacing=list(A_H), pos=p1, val=vel, expect=dict(x=qx),
                    tol=EPS))
    out.append(dict(name="p2g_fcv_quadratic2", src=PT + ":1502-1553", kind="p2g_fcv", alg="quadratic",
                    res=list(A_RES), origin=list(A_O), spacing=list(A_H), pos=p2, val=vel,
                    expect=dict(y=quad(3.0, 4.0)), tol=EPS))
    out.append(dict(name="p2g_fcv_quadratic3", src=PT + ":1555-1604", kind="p2g_fcv", alg="quadratic",
                    res=list(B_RES), origin=list(A_O), spacing=list(A_H), pos=p3, val=vel,
                    expect=dict(z=quad(5.0, 6.0)), tol=EPS))
    out += particle_motion_cases()
    return out


def particle_motion_cases():
    """advectParticlesInVelField, buildFluidMarkers and statisticalFluidDensity
    (UnitTest-ParticlesAdvection.cu:3441-4196).  The advection cases use the
    linear velocity u = 10 i, v = 20 j, w = 30 k on the faces (i, j, k) of a
    4x4x1 grid, so the sampled velocity is 10, 20, 30 times the position
    relative to the origin in cells and the expected positions are the
    reference test's closed forms, evaluated here in FP64."""
    out = []
    o, h = (-10.0, -20.0, -30.0), (10.0, 20.0, 30.0)
    res = (4, 4, 1)
    base = (o[0] + 1.25 * h[0], o[1] + 1.25 * h[1], -15.0)
    pos = []
    for y in range(4):
        for x in range(4):
            pos += [base[0] + x * 0.5 * h[0], base[1] + y * 0.5 * h[1], base[2]]
    vx = [10.0 * i for k in range(1) for j in range(4) for i in range(5)]
    vy = [20.0 * j for k in range(1) for j in range(5) for i in range(4)]
    vz = [30.0 * k for k in range(2) for j in range(4) for i in range(4)]
    scale = (10.0, 20.0, 30.0)

    def rel(p):   # position relative to the origin, in cells
        return [(p[a] - o[a]) / h[a] for a in range(3)]

    def vel(r):
        return [scale[a] * r[a] for a in range(3)]

    def expect(order, steps):
        e = []
        for q in range(16):
            p = pos[3 * q:3 * q + 3]
            for dt in steps:
                r0 = rel(p)
                if order == 1:
                    k = vel(r0)
                elif order == 2:
                    mid = [p[a] + 0.5 * dt * vel(r0)[a] for a in range(3)]
                    k = vel(rel(mid))
                else:
                    k1 = vel(r0)
                    mid = [p[a] + 0.5 * dt * k1[a] for a in range(3)]
                    k2 = vel(rel(mid))
                    tq = [p[a] + 0.75 * dt * k2[a] for a in range(3)]
                    k3 = vel(rel(tq))
                    k = [(2.0 * k1[a] + 3.0 * k2[a] + 4.0 * k3[a]) / 9.0 for a in range(3)]
                p = [p[a] + dt * k[a] for a in range(3)]
            e += p
        return {str(i): v for i, v in enumerate(e)}

    for name, src, order, dt, cfl, steps in [
            ("advect_rk1", ":3445-3537", 1, 0.1, 1.0, [0.1]),
            ("advect_rk1_3substeps", ":3539-3637", 1, 0.4, 1.1, [0.2, 0.1, 0.1]),
            ("advect_rk2", ":3639-3735", 2, 0.1, 1.0, [0.1]),
            ("advect_rk3", ":3737-3838", 3, 0.1, 1.0, [0.1])]:
        out.append(dict(name=name, src=PT + src, kind="advect", res=list(res), origin=list(o), spacing=list(h),
                        pos=pos, u=vx, v=vy, w=vz, dt=dt, cfl=cfl, order=order, expect=expect(order, steps),
                        tol=EPS))
    # buildFluidMarkers: solid everywhere but cells 5, 6, 9, 10, 13, 14; the
    # 4x4 particles (a 2x2 block per cell) lie in cells 5, 6, 9, 10
    solid = [-1.0] * 64
    for c in (5, 6, 9, 10, 13, 14):
        solid[c] = 1.0
    pos = []
    for y in range(4):
        for x in range(4):
            pos += [base[0] + x * 0.5 * h[0], base[1] + y * 0.5 * h[1], base[2]]
    e = {str(i): (1.0 if i in (5, 6, 9, 10) else 0.0 if i in (13, 14) else 2.0) for i in range(64)}
    out.append(dict(name="fluid_markers", src=PT + ":3843-3917", kind="markers", res=[4, 4, 4], origin=list(o),
                    spacing=list(h), solid=solid, pos=pos, expect=e, tol=EPS))
    # statisticalFluidDensity of particles generated in the 8 central cells, 8 each
    eight = [21, 22, 25, 26, 37, 38, 41, 42]
    fluid = [1.0] * 64
    solid = [-1.0] * 64
    for c in eight:
        fluid[c] = -1.0
        solid[c] = 1.0
    e = {str(i): (8.0 if i in eight else 0.0) for i in range(64)}
    out.append(dict(name="fluid_density_particles", src=PT + ":4139-4196", kind="density", res=[4, 4, 4],
                    origin=list(o), spacing=list(h), fluid=fluid, solid=solid, per_cell=8, expect=e, tol=EPS))
    return out


def main():
    cases = dict(pressure=pressure_cases(), semilagrangian=semilagrangian_cases(),
                 external_forces=external_forces_cases(), extrapolation=extrapolation_cases(),
                 boundary=boundary_cases(), cg=cg_cases(), fields=field_cases(),
                 integration=integration_cases(), particles=particle_cases())
    path = os.path.join(HERE, "reference_cases.json.gz")
    with gzip.open(path, "wt", compresslevel=9) as f:
        json.dump(cases, f, separators=(",", ":"))
    print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
