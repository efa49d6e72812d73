"""CPU checks of the C ABI: the library loads, exports every function that
include/dsmc.h declares, the ctypes mirror has the header's struct layouts
(compared with a C probe compiled against the header), and configuration
errors are rejected before any device call."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

from cudadsmc_amd import _abi, configs

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "dsmc.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dsmc_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    lib = C.CDLL(str(_abi.LIB_PATH))
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(_abi.PROTOTYPES), set(names) ^ set(_abi.PROTOTYPES)


def test_abi_version():
    assert _abi.load_library().dsmc_abi_version() == 1


def test_struct_layout_matches_header(tmp_path):
    probe = tmp_path / "probe.c"
    fields = {
        "dsmc_trap": ["kind", "n_atoms", "nc", "philox_stream", "half_width", "B0", "dBdz", "mu",
                      "T", "T_axis", "n_real", "eta_start", "T_cut", "eta_steps"],
        "dsmc_config": ["precision", "pair_mode", "evaporation", "spin", "spin_substeps",
                        "spin_interp", "lpc", "use_graph", "dt", "mass", "sigma", "gs", "seed",
                        "block_size", "obs_capacity"],
        "dsmc_observables": ["ek", "ep", "v2", "e_evap", "n_alive", "n_evap", "collisions", "step"],
        "dsmc_info": ["precision", "n_keys", "block_size", "step", "obs_capacity"],
        "dsmc_kernel_stat": ["name", "launches", "total_ms", "bytes_per_launch"],
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){"]
    for st, fs in fields.items():
        lines.append(f'printf("{st} size %zu\\n", sizeof({st}));')
        for f in fs:
            lines.append(f'printf("{st} {f} %zu\\n", offsetof({st}, {f}));')
    lines.append("return 0;}")
    probe.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c11", str(probe), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    mirror = {"dsmc_trap": _abi.Trap, "dsmc_config": _abi.Config,
              "dsmc_observables": _abi.Observables, "dsmc_info": _abi.Info,
              "dsmc_kernel_stat": _abi.KernelStat}
    for line in filter(None, out):
        st, what, val = line.split()
        cls = mirror[st]
        if what == "size":
            assert C.sizeof(cls) == int(val), (st, C.sizeof(cls), val)
        else:
            assert getattr(cls, what).offset == int(val), (st, what)


def _create(run):
    lib = _abi.load_library()
    traps = (_abi.Trap * len(run.traps))(*[t.to_c() for t in run.traps])
    cfg = run.to_c()
    h = C.c_void_p()
    st = lib.dsmc_create(C.byref(cfg), traps, len(run.traps), C.byref(h))
    return st, lib.dsmc_last_error(None).decode()


@pytest.mark.parametrize("mutate,status,needle", [
    (lambda r: r.with_(dt=0.0), _abi.DSMC_EINVAL, "dt"),
    (lambda r: r.with_(lpc=0), _abi.DSMC_EINVAL, "lpc"),
    (lambda r: r.with_(block_size=100), _abi.DSMC_EINVAL, "block_size"),
    (lambda r: r.with_(spin_loss=True, spin=False), _abi.DSMC_EINVAL, "spin"),
    (lambda r: r.with_(precision="fp32", lpc=2), _abi.DSMC_ENOTSUP, "lpc"),
    (lambda r: r.with_(traps=r.traps * 2), _abi.DSMC_ENOTSUP, "FP64 form runs one trap"),
    (lambda r: r.with_(traps=[configs.Trap(kind="quadrupole", n_atoms=10, dBdz=0.0)]),
     _abi.DSMC_EINVAL, "quadrupole"),
    (lambda r: r.with_(traps=[configs.Trap(n_atoms=10, nc=(0, 4, 4))]), _abi.DSMC_EINVAL, "nc"),
    (lambda r: r.with_(obs_capacity=-1), _abi.DSMC_EINVAL, "obs_capacity"),
    (lambda r: r.with_(mcwf=True), _abi.DSMC_ENOTSUP, "spin-flip"),
    (lambda r: r.with_(mcwf=True, spin=True, precision="fp32"), _abi.DSMC_ENOTSUP, "spin-flip"),
])
def test_invalid_configuration_rejected(mutate, status, needle):
    st, msg = _create(mutate(configs.quadrupole_1e4()))
    assert st == status
    assert needle in msg


def test_invalid_sort_mode_rejected():
    lib = _abi.load_library()
    run = configs.quadrupole_1e4()
    traps = (_abi.Trap * 1)(run.traps[0].to_c())
    cfg = run.to_c()
    cfg.sort_mode = 7
    h = C.c_void_p()
    assert lib.dsmc_create(C.byref(cfg), traps, 1, C.byref(h)) == _abi.DSMC_EINVAL
    assert "sort_mode" in lib.dsmc_last_error(None).decode()


def test_null_arguments_rejected():
    lib = _abi.load_library()
    h = C.c_void_p()
    assert lib.dsmc_create(None, None, 1, C.byref(h)) == _abi.DSMC_EINVAL
    assert lib.dsmc_step(None, 1) == _abi.DSMC_EINVAL
    assert lib.dsmc_destroy(None) == _abi.DSMC_OK
    assert lib.dsmc_sort_by_key(None, None, 10, 8, None) == _abi.DSMC_EINVAL
    assert lib.dsmc_remove_if(None, None, -1, None, None) == _abi.DSMC_EINVAL


def test_product_has_no_oracle_dependency():
    """The shipped library and package never load the oracle."""
    out = subprocess.run(["nm", "-D", "--undefined-only", str(_abi.LIB_PATH)], check=True,
                         capture_output=True, text=True).stdout
    assert "orc_" not in out
    for f in (ROOT / "cudadsmc_amd").rglob("*.py"):
        assert "oracle" not in f.read_text().replace("no oracle", ""), f
