"""The C-ABI library loads on the CPU and exports every function the public headers declare
(no GPU calls)."""
import os
import re
import subprocess

import gdmr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"^typedef[^;]*;", "", text, flags=re.M)
    names = re.findall(r"^[A-Za-z_][\w \*]*?\b(\w+)\s*\([^;{]*\)\s*;", text, flags=re.M)
    return {n for n in names if n not in ("if", "while")}


def test_headers_fully_exported():
    declared = (_declared("gDMR.h") | _declared("gDMRC.h") | _declared("gDMRServe.h") |
                _declared("gDMRKernels.h"))
    assert {"initDMR", "freeDMR", "solverDMR", "impSVDMR", "expSVDMR", "initLS", "solveLS",
            "linOptLS", "impCsrLS", "dmrServerSubmit", "dmrServerWait"} <= declared
    out = subprocess.run(["nm", "-D", "--defined-only", gdmr.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = declared - exported
    assert not missing, f"declared but not exported: {sorted(missing)}"
    assert set(gdmr.HEADER_SYMBOLS) <= exported
    assert declared <= set(gdmr.HEADER_SYMBOLS) | {"dmrOperatorFn"}


def test_reference_symbol_set_and_soname():
    """The reference's 5 solver + 10 linear-system entry points (gDMR.h:13-34, gDMRC.h:6-24), and
    the soname of its link model (Makefile:18-21)."""
    ref = ("initDMR freeDMR solverDMR impSVDMR expSVDMR initLS freeLS impImgLS expImgLS impSrcLS "
           "expSrcLS impOptLS expOptLS solveLS linOptLS").split()
    for s in ref:
        assert hasattr(gdmr.lib, s)
    dyn = subprocess.run(["readelf", "-d", gdmr.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    assert "libgDMR.so.1" in dyn
    prog = os.path.join(gdmr.LIB_DIR, "progDMR")
    assert os.path.exists(prog)
    deps = subprocess.run(["readelf", "-d", prog], capture_output=True, text=True).stdout
    assert "libgDMR.so.1" in deps and "$ORIGIN" in deps


def test_cpu_only_entry_points():
    """Entry points that need no device: errors before init, sizes, footprints."""
    assert gdmr.reduction_blocks(1) == 1
    assert gdmr.reduction_blocks(4096) == 8
    assert gdmr.reduction_blocks(2097152) == 512
    assert gdmr.footprint(2097152, 50, 20) > 8 * 2097152 * 91
    its = gdmr.C.c_int(0)
    r = gdmr.lib.solverDMR(0, None, None, None, 0, 1e-8, 10, gdmr.C.byref(its))
    assert r == 1.0 and gdmr.last_error()[0] == gdmr.DMR_ERR_STATE


def test_server_rejects_bad_arguments_without_device():
    """Argument checks of the serving API come before any device call; with no device visible
    creation fails cleanly with DMR_ERR_HIP instead of starting workers."""
    h = gdmr.C.c_void_p()
    assert gdmr.lib.dmrServerCreate(gdmr.C.byref(h), 0, 0, 0) == gdmr.DMR_ERR_ARG
    assert gdmr.lib.dmrServerCreate(None, 0, 2, 0) == gdmr.DMR_ERR_ARG
    assert gdmr.lib.dmrServerCreate(gdmr.C.byref(h), 0, 2, 0) in (gdmr.DMR_OK, gdmr.DMR_ERR_HIP)
    if h.value:
        gdmr.lib.dmrServerDestroy(h)
    fam = gdmr.C.c_int()
    assert gdmr.lib.dmrServerAddFamily(None, 8, 8, None, None, 4, 2, 0,
                                       gdmr.C.byref(fam)) == gdmr.DMR_ERR_ARG
    rid = gdmr.C.c_longlong()
    assert gdmr.lib.dmrServerSubmit(None, 0, None, 1, None, None, 1e-8, 10,
                                    gdmr.C.byref(rid)) == gdmr.DMR_ERR_ARG
    assert gdmr.lib.dmrServerDrain(None) == gdmr.DMR_ERR_ARG


def test_kernel_entry_points_check_shapes_without_device():
    """The kernel-level entry points reject shapes the kernels do not take before any launch."""
    lib = gdmr.lib
    assert lib.dmrKernelTnWorkspace(1000, 20, 512) == 2 * 32 * 32
    assert lib.dmrKernelTnWorkspace(1000, 20, 100) == 0  # rows per workgroup: multiple of 64
    assert lib.dmrKernelGemmTn(0, None, 10, 4, None, 10, 4, 10, 64, 0, None, None, 4, None,
                               None) == gdmr.DMR_ERR_ARG
    one = gdmr.C.c_void_p(8)  # never dereferenced: every call below fails its shape check
    assert lib.dmrKernelGemmTn(0, one, 10, 400, one, 10, 4, 10, 64, 0, one, one, 400, None,
                               None) == gdmr.DMR_ERR_ARG  # p > 384
    assert lib.dmrKernelGemmTn(0, one, 10, 4, one, 10, 4, 10, 64, 3, one, one, 4, None,
                               None) == gdmr.DMR_ERR_ARG  # variant 3 needs a bf16 basis
    assert lib.dmrKernelGemmNn(1, one, 10, 4, one, None, 10, 0.0, 1.0, one, 10, 20, 10, 100, one,
                               None) == gdmr.DMR_ERR_ARG  # Gram partials need s <= 16
    assert lib.dmrKernelGemmNnPair(one, 10, 4, one, one, one, 10, 17, 10, 100,
                                   None) == gdmr.DMR_ERR_ARG
    assert lib.dmrKernelSpmm(1, 10, 30, one, one, one, one, 10, one, 10, 4, 8,
                             None) == gdmr.DMR_ERR_ARG  # bf16 X takes 104-216
    assert lib.dmrKernelCholQr(one, 33, 1e-12, one, one, 1, None, 0, one,
                               None) == gdmr.DMR_ERR_ARG
    assert lib.dmrKernelGramChol(one, 100, 17, 1e-12, one, one, None, 0, one, one,
                                 None) == gdmr.DMR_ERR_ARG  # s > 16
    assert lib.dmrKernelGramChol(one, 100, 10, 1e-12, one, one, None, 0, one, None,
                                 None) == gdmr.DMR_ERR_ARG  # no work buffer
    hdr = open(os.path.join(ROOT, "include", "gDMRKernels.h")).read()
    m = re.search(r"#define DMR_KGRAMCHOL_WORK_BYTES \((\d+) \+ (\d+) \* (\d+) \* (\d+)\)", hdr)
    a, b, c, e = (int(x) for x in m.groups())
    assert a + b * c * e == gdmr.DMR_KGRAMCHOL_WORK_BYTES
    assert lib.dmrKernelLsqBlock(one, 30, 20, 8, one, 4, one, one, one, one, one, one,
                                 None) == gdmr.DMR_ERR_ARG  # rows != cols + bw
    assert lib.dmrKernelLsqFits(100, 16, 16) == 1 and lib.dmrKernelLsqFits(100, 40, 16) == 0
    gbps = gdmr.C.c_double(0.0)
    assert lib.dmrProbeReadBandwidth(one, 1 << 20, 4, None,
                                     gdmr.C.byref(gbps)) == gdmr.DMR_ERR_ARG  # < 64 MiB
    assert lib.dmrProbeReadBandwidth(one, (64 << 20) + 8, 4, None,
                                     gdmr.C.byref(gbps)) == gdmr.DMR_ERR_ARG  # not 16-byte units
