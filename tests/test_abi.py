"""CPU checks of the C ABI: the library loads, exports every declared symbol,
and reports errors (no compute calls; these run without a GPU)."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    with open(os.path.join(ROOT, "include", "ohpic.h")) as f:
        txt = f.read()
    return sorted(set(re.findall(r"^(?:int|const char \*)\s*(ohpic_\w+)\(", txt, re.M)))


@pytest.fixture(scope="module")
def L():
    from ohhelppic3d_amd import _lib
    if not _lib.available():
        from ohhelppic3d_amd import build
        build.build()
    return _lib


def test_every_declared_symbol_is_exported(L):
    names = _declared()
    assert len(names) >= 25
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (ohpic_\w+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    assert set(names) == set(L.EXPORTS)
    lib = L.lib()
    for n in names:
        assert getattr(lib, n) is not None


def test_library_is_gfx950_code_object(L):
    # the offload bundle embedded in the library names its target
    with open(L.LIB_PATH, "rb") as f:
        blob = f.read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob


def test_config_default_and_footprint(L):
    lib = L.lib()
    c = L.Config()
    assert lib.ohpic_config_default(C.byref(c)) == 0
    assert (c.nx, c.ny, c.nz, c.nspecies) == (32, 32, 32, 2)
    b = C.c_int64()
    assert lib.ohpic_estimate_footprint(C.byref(c), C.byref(b)) == 0
    assert b.value > 2 * 2 * 48 * c.capacity[0]
    # the 256^3 x 64 ppc FP64 workload fits in one MI355X (288 GB)
    c.nx = c.ny = c.nz = 256
    c.capacity[0] = c.capacity[1] = 64 * 256 ** 3
    c.slack = 0.05
    assert lib.ohpic_estimate_footprint(C.byref(c), C.byref(b)) == 0
    assert 200e9 < b.value < 260e9


def test_errors_without_gpu_or_with_bad_config(L):
    lib = L.lib()
    c = L.Config()
    lib.ohpic_config_default(C.byref(c))
    ctx = C.c_void_p()
    c.nspecies = 0
    assert lib.ohpic_create(C.byref(c), C.byref(ctx)) == L.EINVAL
    assert b"nspecies" in lib.ohpic_last_error(None)
    assert lib.ohpic_create(None, C.byref(ctx)) == L.EINVAL
    assert lib.ohpic_config_default(None) == L.EINVAL
    out = (C.c_double * 4)()
    assert lib.ohpic_energy(None, out) == L.EINVAL


def test_host_mirror_validation(L):
    import ohhelppic3d_amd as ohp
    with pytest.raises(ohp.OhpicError):
        ohp.create_particle_mover("leapfrog")
    assert isinstance(ohp.create_particle_mover("boris"), ohp.ParticleMoverBoris)
    q = ohp.charge_per_macro_particle([1.0, 0.1], [-1.0, 0.01], [64, 64])
    assert q[0] == pytest.approx(-1 / 64) and q[1] == pytest.approx(1 / 64)
    with pytest.raises(ohp.OhpicError):
        ohp.SimConfig(qm=[-1.0], q_macro=[1.0, 2.0], capacity=[1]).to_c()


def test_group_and_autotune_errors_without_gpu(L):
    # argument checks come first; with no GPU in this process the group and
    # context constructors report ENODEV instead of touching a device
    lib = L.lib()
    g = C.c_void_p()
    assert lib.ohpic_group_create(0, None) == L.EINVAL
    rc = lib.ohpic_group_create(0, C.byref(g))
    assert rc in (L.ENODEV, L.OK)
    if rc == L.OK:  # a GPU is visible: nothing else to check here
        lib.ohpic_group_destroy(g)
        return
    assert lib.ohpic_group_step(None, 1) == L.EINVAL
    assert lib.ohpic_group_set_batch(None, 4) == L.EINVAL
    assert lib.ohpic_group_attach(None, None) == L.EINVAL
    assert lib.ohpic_group_size(None, None) == L.EINVAL
    assert lib.ohpic_group_destroy(None) == L.OK
    r = L.TuneResult()
    assert lib.ohpic_autotune(None, 1, 1, 1, C.byref(r)) == L.EINVAL
    c = L.Config()
    lib.ohpic_config_default(C.byref(c))
    assert lib.ohpic_autotune(C.byref(c), 0, 1, 1, C.byref(r)) == L.EINVAL
    ctx = C.c_void_p()
    assert lib.ohpic_create(C.byref(c), C.byref(ctx)) == L.ENODEV
    assert b"device" in lib.ohpic_last_error(None).lower()


def test_create_rejects_oversized_axes_and_bad_modes(L):
    lib = L.lib()
    c = L.Config()
    ctx = C.c_void_p()
    for field, value, msg in (("nx", 2048, b"axis lengths"), ("precision", 7, b"precision"),
                              ("weight_order", 2, b"weight_order"), ("chunks", 32768, b"chunks"),
                              ("chunks", -1, b"chunks")):
        lib.ohpic_config_default(C.byref(c))
        setattr(c, field, value)
        assert lib.ohpic_create(C.byref(c), C.byref(ctx)) == L.EINVAL, field
        assert msg in lib.ohpic_last_error(None), field
    lib.ohpic_config_default(C.byref(c))
    c.qm[1] = 0.0
    assert lib.ohpic_create(C.byref(c), C.byref(ctx)) == L.EINVAL
    assert b"qm == 0" in lib.ohpic_last_error(None)


def test_footprint_estimate_grows_with_particles_and_slack(L):
    lib = L.lib()
    c = L.Config()
    lib.ohpic_config_default(C.byref(c))
    b1, b2, b3 = C.c_int64(), C.c_int64(), C.c_int64()
    assert lib.ohpic_estimate_footprint(C.byref(c), C.byref(b1)) == 0
    c.capacity[0] *= 2
    assert lib.ohpic_estimate_footprint(C.byref(c), C.byref(b2)) == 0
    c.slack = 0.5
    assert lib.ohpic_estimate_footprint(C.byref(c), C.byref(b3)) == 0
    assert b1.value < b2.value < b3.value
    # chunks per tile beyond the work item's 16-bit field are refused
    c.chunks = 40000
    assert lib.ohpic_estimate_footprint(C.byref(c), C.byref(b3)) == L.EINVAL
    assert b"chunks" in lib.ohpic_last_error(None)
    c.chunks = 0
    c.precision = L.MIXED
    b4 = C.c_int64()
    assert lib.ohpic_estimate_footprint(C.byref(c), C.byref(b4)) == 0
    assert b4.value < b3.value   # 28 B per particle against FP64's 48 B
