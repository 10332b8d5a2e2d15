"""The host side of the boundary, without a GPU: the C-ABI library loads and exports every function
include/pyhype_hip.h declares, the ctypes descriptors have the C layout, plan creation rejects bad
descriptors before it touches a device, and the host packing of block layouts and geometry
(including the Gauss-point offsets) agrees with the oracle's geometry."""
import ctypes as C
import re
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import pyhype_oracle as O
from pyhype_amd import _native as nat
from pyhype_amd import workloads as W
from pyhype_amd.backend import CaseSpec, layout_cases
from pyhype_amd.mesh import BlockGeometry
from pyhype_amd.solvers import mesh_dict, validate_config

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "pyhype_hip.h"


def _declared():
    txt = HEADER.read_text()
    return sorted(set(re.findall(r"^(?:int|int64_t|const char\*)\s+(phx_\w+)\(", txt, re.M)))


def _defines():
    out = {}
    for name, val in re.findall(r"^#define\s+(PHX_\w+)\s+\(?(-?\d+)\)?", HEADER.read_text(), re.M):
        out[name] = int(val)
    return out


def test_library_loads_and_exports_every_declared_function():
    lib = nat.load()
    names = _declared()
    assert len(names) >= 18
    for n in names:
        assert callable(getattr(lib, n)), n
    assert set(names) == set(nat.EXPORTS)


LLVM_BIN = Path("/opt/rocm/llvm/bin")


@pytest.mark.skipif(not (LLVM_BIN / "llvm-objdump").exists(), reason="ROCm LLVM tools not available")
def test_device_code_makes_no_function_calls(tmp_path):
    """Every kernel of the built gfx950 code object is call-free (no s_swappc_b64): the per-item
    functions are force-inlined, because the outlined HLLL Gauss-face kernels returned NaN."""
    fatbin, code = tmp_path / "fatbin.bin", tmp_path / "gfx950.o"
    subprocess.run([str(LLVM_BIN / "llvm-objcopy"), f"--dump-section=.hip_fatbin={fatbin}", str(nat.LIB_PATH),
                    str(tmp_path / "host.so")], check=True, capture_output=True)
    subprocess.run([str(LLVM_BIN / "clang-offload-bundler"), "--unbundle", "--type=o",
                    "--targets=hipv4-amdgcn-amd-amdhsa--gfx950", f"--input={fatbin}", f"--output={code}"],
                   check=True, capture_output=True)
    dis = subprocess.run([str(LLVM_BIN / "llvm-objdump"), "-d", "--mcpu=gfx950", str(code)], check=True,
                         capture_output=True, text=True).stdout
    assert dis.count("s_endpgm") > 50
    calls = [ln.strip() for ln in dis.splitlines() if "s_swappc" in ln]
    assert not calls, f"{len(calls)} device-function calls, e.g. {calls[:2]}"


def test_python_constants_match_the_header():
    d = _defines()
    assert d["PHX_ABI_VERSION"] == nat.PHX_ABI_VERSION
    assert (d["PHX_OK"], d["PHX_E_INVALID"], d["PHX_E_HIP"], d["PHX_E_UNREALIZABLE"], d["PHX_E_STATE"]) == \
        (nat.PHX_OK, nat.PHX_E_INVALID, nat.PHX_E_HIP, nat.PHX_E_UNREALIZABLE, nat.PHX_E_STATE)
    assert (d["PHX_E"], d["PHX_W"], d["PHX_N"], d["PHX_S"]) == (nat.E, nat.W, nat.N, nat.S)
    assert (d["PHX_SIDE_NEIGHBOR"], d["PHX_SIDE_OPEN"], d["PHX_SIDE_WALL"], d["PHX_SIDE_OUTLET"]) == \
        (nat.SIDE_NEIGHBOR, nat.SIDE_OPEN, nat.SIDE_WALL, nat.SIDE_OUTLET)
    assert (d["PHX_FLUX_ROE"], d["PHX_FLUX_HLLL"]) == (nat.FLUX_ROE, nat.FLUX_HLLL)
    assert (d["PHX_RECON_CONSERVATIVE"], d["PHX_RECON_PRIMITIVE"]) == (nat.RECON_CONSERVATIVE, nat.RECON_PRIMITIVE)
    assert {"Venkatakrishnan": d["PHX_LIMITER_VENKATAKRISHNAN"], "BarthJespersen": d["PHX_LIMITER_BARTH_JESPERSEN"],
            "VanAlbada": d["PHX_LIMITER_VAN_ALBADA"], "VanLeer": d["PHX_LIMITER_VAN_LEER"]} == nat.LIMITERS
    assert (d["PHX_GEOM_GENERAL"], d["PHX_GEOM_CARTESIAN"]) == (nat.GEOM_GENERAL, nat.GEOM_CARTESIAN)
    assert d["PHX_MAX_QUADRATURE_POINTS"] == 3


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_ctypes_descriptors_have_the_c_layout(tmp_path):
    """sizeof/offsetof of phx_block_desc and phx_plan_desc from a C compiler against ctypes."""
    structs = {"phx_block_desc": nat.BlockDesc, "phx_plan_desc": nat.PlanDesc}
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "pyhype_hip.h"', "int main(void) {"]
    for cname, cls in structs.items():
        lines.append(f'printf("%zu\\n", sizeof({cname}));')
        for f in cls._fields_:
            lines.append(f'printf("%zu\\n", offsetof({cname}, {f[0]}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c99", f"-I{HEADER.parent}", str(src), "-o", str(exe)], check=True)
    vals = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    expect = []
    for cls in structs.values():
        expect.append(C.sizeof(cls))
        expect += [getattr(cls, f[0]).offset for f in cls._fields_]
    assert vals == expect


def _desc(nq=1, geom_extra=0, abi=None):
    """A one-block descriptor (Cartesian 4x3, walls) that passes validation up to the device calls."""
    g = BlockGeometry({"SW": [0.0, 0.0], "SE": [4.0, 0.0], "NW": [0.0, 3.0], "NE": [4.0, 3.0]}, 4, 3)
    geom = g.pack_cartesian()
    if nq > 1:
        geom = np.concatenate([geom, g.pack_quadrature(nq)])
    geom = np.ascontiguousarray(np.concatenate([geom, np.zeros(geom_extra)]))
    blk = (nat.BlockDesc * 1)()
    b = blk[0]
    b.nx, b.ny, b.case_id, b.geom_kind = 4, 3, 0, nat.GEOM_CARTESIAN
    for s in range(4):
        b.side_type[s] = nat.SIDE_WALL
        b.neighbor[s] = -1
    b.state_offset = 0
    b.geom = geom.ctypes.data_as(C.POINTER(C.c_double))
    b.geom_len = geom.size
    a = np.array([[1.0]])
    tf = np.array([1.0])
    pd = nat.PlanDesc()
    pd.abi_version = nat.PHX_ABI_VERSION if abi is None else abi
    pd.nblocks, pd.blocks, pd.ncases = 1, blk, 1
    pd.flux, pd.limiter, pd.nstages = nat.FLUX_ROE, 0, 1
    pd.butcher_a = a.ctypes.data_as(C.POINTER(C.c_double))
    pd.gamma, pd.gm1, pd.one_over_gm1, pd.g_over_gm1, pd.cfl = 1.4, 1.4 - 1, 1 / 0.4, 1.4 / 0.4, 0.8
    pd.t_final = tf.ctypes.data_as(C.POINTER(C.c_double))
    pd.reconstruction = nat.RECON_CONSERVATIVE
    pd.quadrature_points = nq
    return pd, (geom, blk, a, tf)


@pytest.mark.parametrize("change,message", [
    (dict(abi=2), "ABI version mismatch"),
    (dict(nq=4), "quadrature_points must be 1, 2 or 3"),
    (dict(nq=0), "quadrature_points must be 1, 2 or 3"),
    (dict(nq=2, geom_extra=5), "geometry length mismatch"),
    (dict(geom_extra=1), "geometry length mismatch"),
])
def test_plan_create_rejects_bad_descriptors_before_any_device_call(change, message):
    lib = nat.load()
    nq = change.get("nq", 1)
    pd, keep = _desc(nq=max(nq, 1) if nq <= 3 else 1, geom_extra=change.get("geom_extra", 0), abi=change.get("abi"))
    pd.quadrature_points = nq
    out = C.c_void_p()
    rc = lib.phx_plan_create(C.byref(pd), C.byref(out))
    assert rc == nat.PHX_E_INVALID
    assert message in lib.phx_last_error().decode()
    assert not out.value


@pytest.mark.parametrize("nq", [1, 2, 3])
def test_layout_offsets_neighbours_and_geometry_lengths(nq):
    cfg, mesh = W.explosion_multiblock(12, 2, 3)
    lay = layout_cases([CaseSpec(mesh_dict(mesh), 12, 12, 1.0)], quadrature_points=nq)
    assert lay.state_len == 6 * 4 * 144
    by_index = {e.index: e for e in lay.blocks}
    opposite = {0: 1, 1: 0, 2: 3, 3: 2}
    for e, d in zip(lay.blocks, lay.descs):
        assert d.state_offset == 4 * 144 * e.index
        base = (8 + 11 * 12 + 11 * 12 + 144) if e.cartesian else (7 * 12 * 13 * 2 + 5 * 144)
        assert d.geom_len == base + (8 * nq * 144 if nq > 1 else 0)
        for s in range(4):
            if d.side_type[s] == nat.SIDE_NEIGHBOR:
                o = by_index[d.neighbor[s]]
                assert o.sides[opposite[s]] == nat.SIDE_NEIGHBOR and o.neighbors[opposite[s]] == e.index


@pytest.mark.parametrize("nq", [1, 2, 3])
@pytest.mark.parametrize("vert", [{"SW": [0.0, 0.0], "SE": [4.0, 0.0], "NW": [0.0, 3.0], "NE": [4.0, 3.0]},
                                  {"SW": [0.0, 0.0], "SE": [5.0, 0.8], "NW": [0.3, 10.0], "NE": [5.6, 9.1]}],
                         ids=["rectangle", "quad"])
def test_packed_geometry_and_gauss_offsets_equal_the_oracle(vert, nq):
    """The kernels' geometry is the oracle's, bit for bit: centroids, areas, face arrays, and the
    Gauss-point offsets xq - xc, yq - yc in the packed order (face E, W, N, S; point; x then y)."""
    nx, ny = 7, 5
    g = BlockGeometry(vert, nx, ny)
    o = O.block_geometry(vert, nx, ny, nq)
    for a, b in ((g.xc, o.xc), (g.yc, o.yc), (g.A, o.A), (g.dx, o.dx), (g.dy, o.dy)):
        assert np.array_equal(a, b)
    for d in O.DIRS:
        assert np.array_equal(g.face[d].L, o.face[d].L)
        assert np.array_equal(g.face[d].theta, o.face[d].theta, equal_nan=True)
    packed = g.pack_quadrature(nq)
    expect = []
    for d in O.DIRS:
        for qx, qy, _ in o.qpts[d]:
            expect += [(qx - o.xc).ravel(), (qy - o.yc).ravel()]
    assert np.array_equal(packed, np.concatenate(expect))
    if nq == 1:   # the 1-point offsets are the midpoint offsets the packed layouts carry
        assert np.array_equal(packed[: nx * ny], (g.face["E"].qx - g.xc).ravel())


def test_config_validation_errors():
    cfg, _ = W.explosion_single(8)
    validate_config(cfg)
    for attr, val, exc in (("fvm_num_quadrature_points", 4, KeyError), ("fvm_slope_limiter_type", "Minmod", ValueError),
                           ("fvm_flux_function_type", "HLLE", ValueError), ("fvm_spatial_order", 1, ValueError),
                           ("use_JIT", False, ValueError)):
        old = getattr(cfg, attr)
        setattr(cfg, attr, val)
        with pytest.raises(exc):
            validate_config(cfg)
        setattr(cfg, attr, old)
    for nq in (2, 3):
        cfg.fvm_num_quadrature_points = nq
        validate_config(cfg)


def test_config_log_ranks_and_printout():
    """SolverConfig (solver_config.py:130-138): show_log_for_procs defaults to rank 0 and keeps a
    list or "all" as given; str() prints one tab-indented "name: value" line per slot."""
    cfg, _ = W.explosion_single(8)
    assert cfg.show_log_for_procs == [0]
    text = str(cfg)
    lines = text.split("\n")
    assert text.endswith("\n") and lines[-1] == "" and len(lines) == len(cfg.__slots__) + 1
    for name, line in zip(cfg.__slots__, lines):
        assert line == f"\t{name}: {getattr(cfg, name)}"
    from pyhype_amd.solver_config import SolverConfig
    import inspect
    args = {k: getattr(cfg, k) for k in inspect.signature(SolverConfig.__init__).parameters
            if k not in ("self", "show_log_for_procs")}
    for v in ("all", [1, 3]):
        assert SolverConfig(**args, show_log_for_procs=v).show_log_for_procs == v


def test_backend_butcher_tables_equal_the_oracle_tables():
    """Every table the backend accepts is the oracle's table, entry for entry (same doubles)."""
    from types import SimpleNamespace

    from pyhype_amd.time_marching import TABLES, butcher_table

    for name in TABLES:
        if name.startswith("Generic"):
            for alpha in (0.3, 0.4, 1.0) if name == "Generic2" else (0.3, 0.4, 0.8):
                got = butcher_table(name, SimpleNamespace(alpha=alpha))
                ref = (O.generic2 if name == "Generic2" else O.generic3)(alpha)
                assert got == ref and all(type(a) is type(b) for r, q in zip(got, ref) for a, b in zip(r, q))
        else:
            assert butcher_table(name) == O.BUTCHER[name], name
    assert butcher_table("SSPRK2") == O.generic2(1)


def test_mesh_dict_recovered_from_block_descriptions():
    """The reference's solver keeps BlockDescription objects only (solvers/base.py:90-100); a binding
    inside the reference's solver recovers the mesh dictionary from their info / geometry.vertices /
    neighbors / bc attributes (mesh_dict_from_block_descriptions).  Stand-ins with the same attribute structure round-trip the 2x4 mesh."""
    from types import SimpleNamespace as NS

    from pyhype_amd.mesh import mesh_dict_from_block_descriptions

    _, gen = W.explosion_multiblock(8)
    m = mesh_dict(gen)
    descs = {}
    for n, d in m.items():
        side = lambda p: NS(E=d[p + "E"], W=d[p + "W"], N=d[p + "N"], S=d[p + "S"])
        bc = side("BCType")
        for c in ("NE", "NW", "SE", "SW"):
            setattr(bc, c, d.get("BCType" + c))
        descs[n] = NS(info=NS(nBLK=d["nBLK"]), geometry=NS(vertices=NS(NE=d["NE"], NW=d["NW"], SE=d["SE"], SW=d["SW"])),
                      neighbors=side("Neighbor"), bc=bc)
    got = mesh_dict_from_block_descriptions(descs)
    for n, d in m.items():
        for k, v in d.items():
            if k.startswith("Neighbor") and len(k) == 10:   # corner neighbours: BlockInfo keeps none
                continue
            assert got[n][k] == v, (n, k)


# Settings of the reference's example configurations (examples/<name>/config.py and mesh.py):
# (nx, ny, flux, integrator, reconstruction, CFL, t_final, a_inf, blocks, domain corners of block 0)
EXAMPLE_SETTINGS = {
    # examples/explosion/config.py:13-31, mesh.py (one block, NW [0, 20], NE [10, 20])
    "explosion": (40, 40, "Roe", "RK4", "ConservativeState", 0.7, 0.07, 343.0, [0], ([0, 0], [10, 20])),
    # examples/explosion_multi/config.py:13-31, mesh.py:3-14 (2 x 4 blocks on [0,10] x [0,20])
    "explosion_multi": (150, 150, "Roe", "RK4", "ConservativeState", 0.7, 0.07, 343.0, list(range(8)),
                        ([0, 0], [5, 5])),
    # examples/implosion/config.py:15-32, mesh.py (one block on [0,10]^2)
    "implosion": (40, 40, "Roe", "RK2", "ConservativeState", 0.4, 0.1, 343.0, [0], ([0, 0], [10, 10])),
    # examples/shockbox/config.py:15-32, mesh.py (one block on [0,10]^2, numbered 1)
    "shockbox": (50, 50, "Roe", "RK2", "ConservativeState", 0.4, 2.0, 343.0, [1], ([0, 0], [10, 10])),
    # examples/dmr/config.py:13-30, mesh.py:1-33 (4 x 1 blocks over the 30 degree ramp)
    "dmr": (50, 50, "HLLL", "RK2", "PrimitiveState", 0.4, 0.25, 343.0, [0, 1, 2, 3], ([0, 0], [1, 2 / np.sqrt(3)])),
}


def test_example_runs_use_the_reference_example_settings():
    """tests/golden/make_examples.py builds each example with the settings of the reference's example
    files: mesh size, flux, integrator, reconstruction, CFL, final time, fluid, block numbering and the
    extent of the first block; Venkatakrishnan, Green-Gauss, one Gauss point everywhere."""
    from pyhype_amd.solvers import mesh_dict
    from tests.golden import make_examples as ME

    assert set(ME.EXAMPLES) <= set(EXAMPLE_SETTINGS)
    for name, (nx, ny, flux, integ, recon, cfl, t_final, a_inf, blocks, (sw, ne)) in EXAMPLE_SETTINGS.items():
        cfg, mesh, cap = ME.example(name)
        assert (cfg.nx, cfg.ny, cfg.fvm_flux_function_type, cfg.time_integrator) == (nx, ny, flux, integ), name
        assert cfg.reconstruction_type.__name__ == recon, name
        assert (cfg.CFL, cfg.t_final, cfg.fluid.far_field.a) == (cfl, t_final, a_inf), name
        assert (cfg.fvm_slope_limiter_type, cfg.fvm_gradient_type, cfg.fvm_num_quadrature_points,
                cfg.fvm_spatial_order) == ("Venkatakrishnan", "GreenGauss", 1, 2), name
        assert cap is None, name
        md = mesh_dict(mesh)
        assert sorted(md) == blocks, name
        b0 = md[blocks[0]]
        assert np.allclose(b0["SW"], sw) and np.allclose(b0["NE"], ne), (name, b0["SW"], b0["NE"])
