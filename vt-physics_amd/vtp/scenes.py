"""Particle generators and the benchmark / parity workloads.

The generators reproduce the reference's lattice rules
(Simulator/Library/Source/Model/ModelHandler.cpp:43-175), including its FP32
accumulation loops and the box's asymmetric wall layers (layerNum layers on
the low side, layerNum+1 on the high side, :130-153), so particle counts and
positions match what the reference would attach.

Seeded jitter uses a counter-based SplitMix64 stream of (seed, particle id,
component), so the same particle gets the same offset whatever the order of
generation.
"""
from __future__ import annotations

import json
import math
from dataclasses import dataclass, field

import numpy as np

from ._abi import MAT_BOUNDARY, MAT_FLUID

f32 = np.float32


# --------------------------------------------------------------- generators
def _accumulate(start: float, stop: float, step: float) -> np.ndarray:
    """`for (float v = start; v < stop; v += step)` in FP32 (ModelHandler.cpp:50-64)."""
    out = []
    v = f32(start)
    stop = f32(stop)
    step = f32(step)
    while v < stop:
        out.append(v)
        v = f32(v + step)
    return np.array(out, dtype=np.float32)


def particle_cube(radius: float, lb, size) -> np.ndarray:
    """ModelHandler::generateParticleCubeElements: z outer, y, x inner."""
    r = f32(radius)
    d = f32(2 * r)
    lb = [f32(v) for v in lb]
    size = [f32(v) for v in size]
    xs = _accumulate(f32(r + lb[0]), f32(lb[0] + size[0]), d)
    ys = _accumulate(f32(r + lb[1]), f32(lb[1] + size[1]), d)
    zs = _accumulate(f32(r + lb[2]), f32(lb[2] + size[2]), d)
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1).astype(np.float32)


def particle_box(radius: float, lb, size, layer_num: int) -> np.ndarray:
    """ModelHandler::generateParticleBoxElements: i (x) outer, j, k (z) inner."""
    r = f32(radius)
    lb = [f32(v) for v in lb]
    size = [f32(v) for v in size]
    num = [int(float(f32(size[a] + r)) / (2.0 * float(r))) for a in range(3)]
    i = np.arange(num[0] + 1)
    j = np.arange(num[1] + 1)
    k = np.arange(num[2] + 1)
    I, J, K = np.meshgrid(i, j, k, indexing="ij")
    I, J, K = I.ravel(), J.ravel(), K.ravel()
    L = layer_num
    keep = ((I < L) | (I >= num[0] - L) | (J < L) | (J >= num[1] - L) | (K < L) | (K >= num[2] - L))
    I, J, K = I[keep], J[keep], K[keep]
    base = [float(f32(lb[a] + r)) for a in range(3)]
    x = (base[0] + 2.0 * float(r) * I).astype(np.float32)
    y = (base[1] + 2.0 * float(r) * J).astype(np.float32)
    z = (base[2] + 2.0 * float(r) * K).astype(np.float32)
    return np.stack([x, y, z], axis=1)


def particle_sphere(radius: float, center, volume_radius: float) -> np.ndarray:
    """ModelHandler::generateParticleSphereElements."""
    gap = f32(radius * 2.0)
    nps = int(math.ceil(f32(volume_radius) / gap))
    out = []
    c = [f32(v) for v in center]
    vr = f32(volume_radius)
    for i in range(-nps, nps + 1):
        for j in range(-nps, nps + 1):
            for k in range(-nps, nps + 1):
                p = (f32(f32(i) * gap + c[0]), f32(f32(j) * gap + c[1]), f32(f32(k) * gap + c[2]))
                d = (p[0] - c[0]) * (p[0] - c[0]) + (p[1] - c[1]) * (p[1] - c[1]) + (p[2] - c[2]) * (p[2] - c[2])
                if f32(d) <= f32(vr * vr):
                    out.append(p)
    return np.array(out, dtype=np.float32).reshape(-1, 3)


def particle_cylinder(radius: float, center, bottom_radius: float, height: float) -> np.ndarray:
    """ModelHandler::generateParticleCylinderElements."""
    r = f32(radius)
    d = f32(2 * r)
    R = f32(bottom_radius)
    top = (f32(center[0]), f32(f32(center[1]) + f32(height) / f32(2)), f32(center[2]))
    out = []
    y = f32(top[1] - r)
    while y >= f32(top[1] - f32(height)):
        x = f32(f32(top[0] - R) + r)
        while x <= f32(top[0] + R):
            m_cos = f32(abs(top[0] - x) / R)
            length = f32(R * f32(math.sqrt(float(f32(1 - m_cos * m_cos)))))
            z = f32(f32(top[2] - length) + r)
            while z <= f32(top[2] + length):
                out.append((x, y, z))
                z = f32(z + d)
            x = f32(x + d)
        y = f32(y - d)
    return np.array(out, dtype=np.float32).reshape(-1, 3)


def particle_plane(radius: float, lb, size, layer_num: int) -> np.ndarray:
    """ModelHandler::generateParticlePlaneElements: z outer, layers in y, x inner."""
    r = f32(radius)
    d = f32(2 * r)
    lb = [f32(v) for v in lb]
    size = [f32(v) for v in size]
    xs = _accumulate(f32(r + lb[0]), f32(lb[0] + size[0]), d)
    zs = _accumulate(f32(r + lb[2]), f32(lb[2] + size[2]), d)
    ys = []
    y = f32(r + lb[1])
    for _ in range(layer_num):
        ys.append(y)
        y = f32(y + d)
    Z, Y, X = np.meshgrid(zs, np.array(ys, np.float32), xs, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1).astype(np.float32)


def generate(config: dict) -> np.ndarray:
    """ModelHandler::generateObjectElements dispatch on genType / objTransferType."""
    t = int(config.get("genType", config.get("objTransferType", -1)))
    r = config["particleRadius"]
    if t == 1:
        return particle_sphere(r, config["volumeCenter"], config["volumeRadius"])
    if t == 2:
        return particle_cube(r, config["lb"], config["size"])
    if t == 3:
        return particle_cylinder(r, config["center"], config["bottomAreaRadius"], config["height"])
    if t == 4:
        return particle_plane(r, config["lb"], config["size"], config["layerNum"])
    if t == 5:
        return particle_box(r, config["lb"], config["size"], config["layerNum"])
    raise ValueError(f"unsupported generation type {t}")


# --------------------------------------------------------------- counter RNG
_M64 = (1 << 64) - 1


def _splitmix64(x: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, ids: np.ndarray, comp: int) -> np.ndarray:
    """U[0,1) from (seed, particle id, component), order-independent."""
    with np.errstate(over="ignore"):
        key = (np.uint64(seed & _M64) * np.uint64(0x100000001B3)) ^ (ids.astype(np.uint64) * np.uint64(8) +
                                                                      np.uint64(comp))
    return (_splitmix64(key) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def jitter(pos: np.ndarray, spacing: float, seed: int, frac: float = 0.1, id_offset: int = 0) -> np.ndarray:
    """pos + U(-frac*spacing, +frac*spacing) per component."""
    ids = np.arange(pos.shape[0], dtype=np.uint64) + np.uint64(id_offset)
    out = pos.astype(np.float64).copy()
    for a in range(3):
        out[:, a] += (2.0 * uniform(seed, ids, a) - 1.0) * frac * spacing
    return out.astype(np.float32)


def normal(seed: int, n: int, sigma: float, stream: int = 7) -> np.ndarray:
    """N(0, sigma^2) per component by Box-Muller on the same counter stream."""
    ids = np.arange(n, dtype=np.uint64)
    out = np.empty((n, 3))
    for a in range(3):
        u1 = uniform(seed ^ (stream << 32), ids, 2 * a)
        u2 = uniform(seed ^ (stream << 32), ids, 2 * a + 1)
        out[:, a] = np.sqrt(-2.0 * np.log(np.maximum(u1, 1e-300))) * np.cos(2 * np.pi * u2)
    return (out * sigma).astype(np.float32)


# --------------------------------------------------------------- workloads
@dataclass
class Obj:
    mat: int
    pos: np.ndarray
    vel: np.ndarray | None = None
    vol_frac: np.ndarray | None = None
    scene: int = 0


@dataclass
class Workload:
    name: str
    solver: str
    config: dict
    objects: list = field(default_factory=list)
    steps: int = 1

    @property
    def n(self) -> int:
        return sum(o.pos.shape[0] for o in self.objects)

    @property
    def n_fluid(self) -> int:
        return sum(o.pos.shape[0] for o in self.objects if o.mat == MAT_FLUID)

    def config_json(self) -> str:
        return json.dumps(self.config)

    def scenes(self) -> int:
        return max(o.scene for o in self.objects) + 1

    def concat(self):
        """Attach-order arrays: pos, vel, vol_frac|None, scene, objects ranges."""
        pos = np.concatenate([o.pos for o in self.objects]).astype(np.float32)
        vel = np.concatenate([o.vel if o.vel is not None else np.zeros_like(o.pos) for o in self.objects])
        P = self.config.get(self.solver, {}).get("Required", {}).get("phaseRestDensity")
        frac = None
        if P is not None:
            frac = np.concatenate([o.vol_frac if o.vol_frac is not None else np.zeros((o.pos.shape[0], len(P)),
                                                                                        np.float32)
                                   for o in self.objects]).astype(np.float32)
        scene = np.concatenate([np.full(o.pos.shape[0], o.scene, np.uint32) for o in self.objects])
        starts, ends, mats = [], [], []
        s = 0
        for o in self.objects:
            starts.append(s)
            s += o.pos.shape[0]
            ends.append(s)
            mats.append(o.mat)
        return pos, vel.astype(np.float32), frac, scene, (np.array(starts, np.uint32), np.array(ends, np.uint32),
                                                          np.array(mats, np.int32))


def _export_section():
    return {"Common": {"exportTargetDir": "out", "exportFilePrefix": "obj", "exportFileType": "PLY"},
            "SolverRequired": {"enable": False, "exportFps": 35, "exportGroupPolicy": "MERGE"}}


def pbf_config(radius=0.01, lb=(-0.1, -0.1, -0.1), size=(1.2, 1.2, 1.2), dt=0.004, iterations=4, xsph=0.01,
               max_nbr=60, rho_b=1000.0):
    return {"PBF": {"Required": {"animationTime": 2.0, "timeStep": dt, "particleRadius": radius,
                                 "simSpaceLB": list(lb), "simSpaceSize": list(size), "maxNeighborNum": max_nbr,
                                 "iterationNum": iterations, "XSPH_k": xsph, "fPartRestDensity": 1000.0,
                                 "bPartRestDensity": rho_b},
                    "Optional": {"enable": False, "gravity": [0, -9.8, 0]}},
            "EXPORT": _export_section()}


def dfsph_config(radius=0.005, lb=(-0.05, -0.05, -0.05), size=(2.1, 1.1, 1.1), dt=0.002, vis=1e-3,
                 div_thr=0.1, inc_thr=1e-3, tension=0.0, max_nbr=60, cfl=False):
    cfg = {"DFSPH": {"Required": {"animationTime": 2.0, "timeStep": dt, "particleRadius": radius,
                                  "simSpaceLB": list(lb), "simSpaceSize": list(size), "maxNeighborNum": max_nbr,
                                  "fPartRestDensity": 1000.0, "bPartRestDensity": 1000.0,
                                  "fPartRestViscosity": vis, "divFreeThreshold": div_thr,
                                  "incompThreshold": inc_thr, "surfaceTensionCoefficient": tension},
                     "Optional": {"enable": False, "gravity": [0, -9.8, 0]}},
           "EXPORT": _export_section()}
    if cfl:
        cfg["DFSPH"]["Extensions"] = {"cflEnable": True, "cflNumber": 0.4, "dtMax": dt}
    return cfg


def immct_config(radius=0.005, lb=(-0.05, -0.05, -0.05), size=(2.1, 2.1, 2.1), dt=0.001, three_phase=True,
                 solver="IMMCT", max_nbr=60):
    if solver == "IMM":
        return {"IMM": {"Required": {"animationTime": 1.0, "timeStep": dt, "particleRadius": radius,
                                     "simSpaceLB": list(lb), "simSpaceSize": list(size), "maxNeighborNum": max_nbr,
                                     "divFreeThreshold": 0.1, "incompThreshold": 1e-3,
                                     "surfaceTensionCoefficient": 0.0, "diffusionCoefficientCf": 0.1,
                                     "momentumExchangeCoefficientCd": 0.5,
                                     "phaseRestDensity": [1000.0, 800.0, 1200.0],
                                     "phaseRestViscosity": [1e-3, 5e-3, 1e-2],
                                     "phaseRestColor": [[255, 0, 0], [0, 255, 0], [0, 0, 255]]},
                        "Optional": {"enable": False, "gravity": [0, -9.8, 0]}},
                "EXPORT": _export_section()}
    rho = [1000.0, 800.0, 1200.0] if three_phase else [1000.0, 800.0]
    col = [[255, 0, 0], [0, 255, 0], [0, 0, 255]][:len(rho)]
    cfg = {"IMMCT": {"Required": {"animationTime": 1.0, "timeStep": dt, "particleRadius": radius,
                                  "simSpaceLB": list(lb), "simSpaceSize": list(size), "maxNeighborNum": max_nbr,
                                  "divFreeThreshold": 0.1, "incompThreshold": 1e-3,
                                  "surfaceTensionCoefficient": 0.0, "diffusionCoefficientCf": 0.1,
                                  "momentumExchangeCoefficientCd0": 0.5, "solventViscosity": 1e-3,
                                  "phaseRestDensity": rho, "phaseRestColor": col,
                                  # x1000 inside the solver (IMMCTCudaImpl.cu:58-61): polymer viscosity 5.0
                                  "solutionBasicViscosity": 0.005, "solutionMaxViscosity": 0.01,
                                  "relaxationTime": 0.05, "shearThinningBasicFactor": 0.5,
                                  "rheologicalThreshold": 0.4},
                     "Optional": {"enable": False, "gravity": [0, -9.8, 0]}},
           "EXPORT": _export_section()}
    if three_phase:
        cfg["IMMCT"]["Extensions"] = {"polymerPhase": 1, "phaseRestViscosity": [1e-3, 1e-3, 1e-2]}
    return cfg


def dam_break(name: str, solver: str, config: dict, n_side, spacing: float, box_size, fluid_lb,
              seed: int = 0, layers: int = 2, jitter_frac: float = 0.1, vel=None, scene: int = 0,
              id_offset: int = 0, phases=None) -> list:
    """Fluid lattice block (n_side = (nx, ny, nz)) + 2-layer box walls at [0, box_size]."""
    r = spacing / 2
    nx, ny, nz = n_side
    fluid = particle_cube(r, fluid_lb, (nx * spacing, ny * spacing, nz * spacing))
    assert fluid.shape[0] == nx * ny * nz, (fluid.shape, n_side)
    if jitter_frac:
        fluid = jitter(fluid, spacing, seed, jitter_frac, id_offset)
    walls = particle_box(r, (0.0, 0.0, 0.0), box_size, layers)
    objs = []
    if phases is None:
        objs.append(Obj(MAT_FLUID, fluid, vel, None, scene))
    else:
        objs.extend(phases(fluid, scene))
    objs.append(Obj(MAT_BOUNDARY, walls, None, None, scene))
    return objs


def cfg0_pbf(seed: int = 0, n_side=(32, 32, 32)) -> Workload:
    """PBF dam break: 32^3 fluid at spacing 0.02 in a [0,1]^3 box, config-0 physics."""
    cfg = pbf_config()
    objs = dam_break("cfg0", "PBF", cfg, n_side, 0.02, (1.0, 1.0, 1.0), (0.04, 0.04, 0.04), seed)
    return Workload("pbf_dam_break_32k", "PBF", cfg, objs, steps=500)


def pbf_1m(seed: int = 0) -> Workload:
    """PBF dam break at 1,000,000 fluid particles: 100^3 at spacing 0.02, box [0,4]x[0,2.5]x[0,2.5].

    The reference's mass-free constraint gradient corrects density errors
    weakly at this particle mass, so the column sinks through the two-layer
    floor within tens of ticks; the simulation space reaches 1.2 m below and
    0.6 m around the box so those particles stay on the grid and in the work.
    """
    cfg = pbf_config(lb=(-0.6, -1.2, -0.6), size=(5.2, 4.0, 3.7))
    objs = dam_break("pbf1m", "PBF", cfg, (100, 100, 100), 0.02, (4.0, 2.5, 2.5), (0.04, 0.04, 0.04), seed)
    return Workload("pbf_dam_break_1m", "PBF", cfg, objs, steps=20)


def cfg1_dfsph(seed: int = 0, n_side=(64, 64, 64), cfl=False) -> Workload:
    """DFSPH dam break: 64^3 fluid at spacing 0.01 in a [0,2]x[0,1]x[0,1] box."""
    cfg = dfsph_config(cfl=cfl)
    objs = dam_break("cfg1", "DFSPH", cfg, n_side, 0.01, (2.0, 1.0, 1.0), (0.02, 0.02, 0.02), seed)
    return Workload("dfsph_dam_break_262k", "DFSPH", cfg, objs, steps=1000)


def dfsph_1m(seed: int = 0) -> Workload:
    """DFSPH dam break at 1,000,000 fluid particles: 100^3 at spacing 0.01, box [0,2]x[0,1.25]x[0,1.25]."""
    cfg = dfsph_config(size=(2.1, 1.35, 1.35))
    objs = dam_break("dfsph1m", "DFSPH", cfg, (100, 100, 100), 0.01, (2.0, 1.25, 1.25), (0.02, 0.02, 0.02), seed)
    return Workload("dfsph_dam_break_1m", "DFSPH", cfg, objs, steps=20)


def three_slabs(P: int):
    """Fluid split into three stacked y-slabs with one-hot phase fractions."""
    def make(fluid: np.ndarray, scene: int):
        y = fluid[:, 1]
        lo, hi = float(y.min()), float(y.max())
        cut1 = lo + (hi - lo) / 3.0
        cut2 = lo + 2.0 * (hi - lo) / 3.0
        slab = np.where(y < cut1, 0, np.where(y < cut2, 1, 2))
        objs = []
        for k in range(3):
            sel = slab == k
            frac = np.zeros((int(sel.sum()), P), np.float32)
            frac[:, min(k, P - 1)] = 1.0
            objs.append(Obj(MAT_FLUID, fluid[sel], None, frac, scene))
        return objs
    return make


def cfg2_immct(seed: int = 0, n_side=(100, 100, 100), box=(2.0, 2.0, 2.0), solver="IMMCT") -> Workload:
    """IMM-CT three-phase mixture: 100^3 fluid at spacing 0.01 in a [0,2]^3 box."""
    cfg = immct_config(size=tuple(b + 0.1 for b in box), solver=solver)
    P = 3
    objs = dam_break("cfg2", solver, cfg, n_side, 0.01, box, (0.02, 0.02, 0.02), seed, phases=three_slabs(P))
    for o in objs:
        if o.mat == MAT_BOUNDARY:
            o.vol_frac = np.zeros((o.pos.shape[0], P), np.float32)
    return Workload("immct_3phase_1m", solver, cfg, objs, steps=300)


def batched_scene_sizes(seed: int = 0, scenes: int = 64, lo: int = 4096, hi: int = 262144) -> np.ndarray:
    u = uniform(seed, np.arange(scenes, dtype=np.uint64), 5)
    return np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo))).astype(np.int64)


def _lattice_dims(n: int) -> tuple:
    side = int(math.ceil(n ** (1.0 / 3.0)))
    return side, side, int(math.ceil(n / (side * side)))


def cfg3_batched(seed: int = 0, scenes: int = 64, lo: int = 4096, hi: int = 262144) -> Workload:
    """64 independent PBF dam breaks with log-uniform fluid counts in [4096, 262144].

    Each scene's fluid is a lattice block of exactly N_f particles (full x-z
    layers, the top layer partial) at spacing 0.02 in its own box.  A
    [0,1]^3 box holds at most 46^3 particles at that spacing, so scenes whose
    block does not fit get a larger cubic box (edge up to 2.0); every scene
    shares one grid geometry sized for the largest box.
    """
    sizes = batched_scene_sizes(seed, scenes, lo, hi)
    s = 0.02
    objs = []
    max_edge = 1.0
    for sc, nf in enumerate(sizes):
        nx, nz, ny = _lattice_dims(int(nf))
        edge = max(1.0, math.ceil((max(nx, ny, nz) * s + 0.12) * 5) / 5)
        max_edge = max(max_edge, edge)
        full = particle_cube(s / 2, (0.04, 0.04, 0.04), (nx * s, ny * s, nz * s))
        order = np.lexsort((full[:, 0], full[:, 2], full[:, 1]))   # fill from the bottom y layer up
        fluid = jitter(full[order][: int(nf)], s, seed + 1000 + sc)
        objs.append(Obj(MAT_FLUID, fluid, None, None, sc))
        objs.append(Obj(MAT_BOUNDARY, particle_box(s / 2, (0, 0, 0), (edge, edge, edge), 2), None, None, sc))
    cfg = pbf_config(lb=(-0.1, -0.1, -0.1), size=(max_edge + 0.2,) * 3)
    return Workload("pbf_batched_64", "PBF", cfg, objs, steps=20)


def cfg4_adjoint(seed: int = 0, n_side=(64, 32, 32)) -> Workload:
    """Differentiable PBF rollout: 64x32x32 lattice at spacing 0.02 with N(0, 0.1^2) initial velocities.

    A 64-particle row at spacing 0.02 spans 1.28 m, so the box is stretched to
    [0,1.6] along x to hold the block; y and z keep [0,1].

    maxNeighborNum is 128 instead of the template's 60: with 60 the lists of
    the floor layer reach the cap within the 100 ticks (6 % of the particles
    per tick), and the reverse passes need complete, symmetric lists
    (j in N(i) <=> i in N(j)), which a cut list is not.
    """
    cfg = pbf_config(size=(1.8, 1.2, 1.2), max_nbr=128)
    n = n_side[0] * n_side[1] * n_side[2]
    vel = normal(seed, n, 0.1)
    objs = dam_break("cfg4", "PBF", cfg, n_side, 0.02, (1.6, 1.0, 1.0), (0.04, 0.04, 0.04), seed, vel=vel)
    return Workload("pbf_adjoint_65k", "PBF", cfg, objs, steps=100)


WORKLOADS = {"cfg0": cfg0_pbf, "pbf1m": pbf_1m, "cfg1": cfg1_dfsph, "dfsph1m": dfsph_1m, "cfg2": cfg2_immct,
             "cfg3": cfg3_batched, "cfg4": cfg4_adjoint}


def from_scene_document(doc: dict, name: str = "scene") -> Workload:
    """A reference scene file (solver config + OBJECTS) as a Workload.

    Follows ObjectManager::createObjectsFromJson (ObjectManager.cpp:22-41)
    and the attach step of the solvers (PBFSolver.cpp:169-221): each object
    is generated from its ObjectComponentConfig (objTransferType selects the
    generator), gets the material epmMaterial, the velocity velocityStart and,
    for IMM / IMM-CT, the fractions phaseFraction of its
    SolverObjectComponentConfig, and is attached in file order.
    """
    solver = next(k for k in doc if k not in ("EXPORT", "OBJECTS"))
    config = {k: v for k, v in doc.items() if k != "OBJECTS"}
    objs = []
    for o in doc.get("OBJECTS", []):
        comp = dict(o["ObjectComponentConfig"], genType=o["objTransferType"])
        pos = generate(comp).astype(np.float32)
        sc = o["SolverObjectComponentConfig"]
        vel = np.tile(np.asarray(sc.get("velocityStart", [0, 0, 0]), np.float32), (pos.shape[0], 1))
        frac = None
        if "phaseFraction" in sc:
            frac = np.tile(np.asarray(sc["phaseFraction"], np.float32), (pos.shape[0], 1))
        objs.append(Obj(int(comp["epmMaterial"]), pos, vel, frac, 0))
    return Workload(name, solver, config, objs, steps=1)
