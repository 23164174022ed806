"""Host-side checks of libcubbs_bbs_hip.so that need no GPU: the library
loads, exports every symbol its headers declare, keeps the reference's ABI
struct layouts (crates/cubbs-bbs/src/gpu_abi.rs mirrors them byte for byte),
and rejects malformed calls with the reference's call-level codes before any
device work (native/cuda/src/cubbs_bbs_cuda.cu:21-23,356-358,400-410)."""
import ctypes
import json
import os
import re
import subprocess
import sys

import pytest

from conftest import ROOT

sys.path.insert(0, os.path.join(ROOT, "cubbs-cred_amd"))
import cubbs_hip  # noqa: E402


def header_functions():
    names = []
    for h in ("cubbs_bbs_gpu_abi.h", "cubbs_hip_ext.h", "cubbs_proofvalue.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b(cubbs_\w+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = cubbs_hip.lib()
    declared = header_functions()
    assert len(declared) >= 20
    missing = [n for n in declared if not hasattr(lib, n)]
    assert not missing, missing
    # the binding's symbol lists agree with the headers
    assert set(cubbs_hip.ABI_SYMBOLS + cubbs_hip.EXT_SYMBOLS) <= set(declared)


def test_loader_mandatory_symbols_present():
    # crates/cubbs-bbs/src/gpu_runtime.rs:303-327 refuses a library without these
    lib = cubbs_hip.lib()
    for name in ("cubbs_bbs_gpu_sign_batch_v2", "cubbs_bbs_gpu_verify_batch_v2",
                 "cubbs_bbs_gpu_proof_gen_batch_v2", "cubbs_bbs_gpu_proof_verify_batch_v2"):
        assert hasattr(lib, name)


def test_struct_layouts_lp64():
    c = cubbs_hip
    assert ctypes.sizeof(c.ByteSlice) == 16
    assert ctypes.sizeof(c.ByteBuffer) == 24
    assert ctypes.sizeof(c.Result) == 16
    assert ctypes.sizeof(c.VerifyItem) == 16 * 3 + 16
    assert ctypes.sizeof(c.SignItem) == 16 * 3 + 16 + 24
    assert ctypes.sizeof(c.ProofGenItem) == 16 * 4 + 16 + 16 + 24
    assert ctypes.sizeof(c.ProofVerifyItem) == 16 * 4 + 16 + 16 + 8
    assert ctypes.sizeof(c.MessagesToScalarItem) == 16 * 2 + 24


def test_header_layouts_compile_and_match():
    # compile a C probe against include/cubbs_bbs_gpu_abi.h and compare sizes/offsets
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "cubbs_bbs_gpu_abi.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(CuBbsBbsGpuResult),
         offsetof(CuBbsBbsGpuResult, bytes_written), sizeof(CuBbsBbsGpuVerifyItem),
         sizeof(CuBbsBbsGpuProofVerifyItem), offsetof(CuBbsBbsGpuProofVerifyItem, total_message_count),
         sizeof(CuBbsBbsGpuProofGenT2PrepareItem), sizeof(CuBbsBbsGpuG1MsmCommonPointsBatchItem),
         sizeof(CuBbsBbsGpuBackendInfo), sizeof(CuBbsBbsGpuSignEFromResidentScalarsItem));
  return 0;
}
'''
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        cpath = os.path.join(d, "probe.c")
        open(cpath, "w").write(src)
        exe = os.path.join(d, "probe")
        subprocess.check_call(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), cpath, "-o", exe])
        out = subprocess.check_output([exe]).decode().split()
    assert [int(v) for v in out] == [16, 8, 64, 104, 96, 16 * 10 + 24 * 4, 56, 8, 16 * 2 + 8 * 3 + 24 * 2]


def test_call_level_validation_without_device():
    lib = cubbs_hip.lib()
    res = (cubbs_hip.Result * 2)()
    items = (cubbs_hip.VerifyItem * 2)()
    # wrong ABI version
    assert lib.cubbs_bbs_gpu_verify_batch_v2(1, items, ctypes.c_size_t(2), res) == cubbs_hip.ERR_INVALID_INPUT
    # empty batch is a success
    assert lib.cubbs_bbs_gpu_verify_batch_v2(2, None, ctypes.c_size_t(0), None) == cubbs_hip.SUCCESS
    # NULL arrays with n > 0
    assert lib.cubbs_bbs_gpu_verify_batch_v2(2, None, ctypes.c_size_t(2), res) == cubbs_hip.ERR_INVALID_INPUT
    assert lib.cubbs_bbs_gpu_sign_batch_v2(2, None, ctypes.c_size_t(1), res) == cubbs_hip.ERR_INVALID_INPUT
    assert lib.cubbs_bbs_gpu_proof_verify_batch_v2(2, None, ctypes.c_size_t(1), res) == cubbs_hip.ERR_INVALID_INPUT
    assert lib.cubbs_bbs_gpu_proof_gen_batch_v2(3, None, ctypes.c_size_t(0), res) == cubbs_hip.ERR_INVALID_INPUT


def test_messages_to_scalar_rejects_mixed_dst_before_device():
    # reference cubbs_bbs_cuda.cu:400-410: one DST per call
    keep = cubbs_hip._Keep()
    items = (cubbs_hip.MessagesToScalarItem * 2)()
    for i, dst in enumerate([b"dst-one", b"dst-two"]):
        items[i].message = keep.slice(b"m")
        items[i].dst = keep.slice(dst)
        _, bb = keep.out(32)
        items[i].scalar_out = bb
    res = (cubbs_hip.Result * 2)()
    rc = cubbs_hip.lib().cubbs_bbs_gpu_messages_to_scalar_batch_v2(2, items, ctypes.c_size_t(2), res)
    assert rc == cubbs_hip.ERR_INVALID_INPUT
    # an output buffer below 32 bytes is rejected too
    items[1].dst = items[0].dst
    items[1].scalar_out = cubbs_hip.ByteBuffer(items[1].scalar_out.data, 31, 0)
    rc = cubbs_hip.lib().cubbs_bbs_gpu_messages_to_scalar_batch_v2(2, items, ctypes.c_size_t(2), res)
    assert rc == cubbs_hip.ERR_INVALID_INPUT


def test_constants_header_is_derived():
    # bls12_381_consts.hpp must equal what gen_consts.py derives from x
    import importlib.util
    import tempfile
    path = os.path.join(ROOT, "cubbs-cred_amd", "csrc", "gen_consts.py")
    spec = importlib.util.spec_from_file_location("gen_consts", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    committed = open(os.path.join(ROOT, "cubbs-cred_amd", "csrc", "bls12_381_consts.hpp")).read()
    with tempfile.TemporaryDirectory() as d:
        mod.HERE = d
        import shutil
        shutil.copy(os.path.join(ROOT, "cubbs-cred_amd", "csrc", "h2c_iso_consts.h"), d)
        mod.main()
        assert open(os.path.join(d, "bls12_381_consts.hpp")).read() == committed


def test_isogeny_headers_identical():
    a = open(os.path.join(ROOT, "oracle", "h2c_iso_consts.h")).read()
    b = open(os.path.join(ROOT, "cubbs-cred_amd", "csrc", "h2c_iso_consts.h")).read()
    assert a == b


def test_isogeny_derivation_reproduces_fixture_generators():
    # runs the derivation end to end (E' order check, Velu, fixture selection)
    out = subprocess.run([sys.executable, os.path.join(ROOT, "oracle", "derive_h2c_iso.py"),
                          "--out-oracle", "/dev/null", "--out-product", "/dev/null"],
                         capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    assert "reproduces Q1 and H_1..H_10" in out.stderr


def test_tune_table_load_parses_and_rejects():
    lib = cubbs_hip.lib()
    ok = b"pippenger log2_terms=16 c=13 S=16 ms=5.1\ncomb log2_terms=18 chunk=4 ms=2.0\n"
    assert lib.cubbs_hip_tune_load(ok) == 0
    assert lib.cubbs_hip_tune_load(b"pippenger log2_terms=16 c=3 S=16\n") == 1   # c out of range
    assert lib.cubbs_hip_tune_load(b"window 5\n") == 1
    assert lib.cubbs_hip_tune_load(b"") == 0   # clears the table


def test_pipeline_tuning_lines_parse_and_reject():
    lib = cubbs_hip.lib()
    ok = (b"pipeline kind=verify log2_items=13 chunk=4 group=8\n"
          b"pipeline kind=proof_verify log2_items=11 chunk=8 group=4\n")
    assert lib.cubbs_hip_tune_load(ok) == 0
    assert lib.cubbs_hip_tune_load(b"pipeline kind=sign log2_items=13 chunk=4 group=8\n") == 1
    assert lib.cubbs_hip_tune_load(b"pipeline kind=verify log2_items=13 chunk=0 group=8\n") == 1
    assert lib.cubbs_hip_tune_load(b"pipeline kind=verify log2_items=13 chunk=4 group=9\n") == 1
    assert lib.cubbs_hip_tune_load(b"pipeline kind=verify log2_items=41 chunk=4 group=8\n") == 1
    assert lib.cubbs_hip_tune_load(b"pipeline kind=verify log2_items=13 chunk=4 group=8 super=4\n") == 0
    assert lib.cubbs_hip_tune_load(b"pipeline kind=verify log2_items=13 chunk=4 group=8 super=0\n") == 0
    assert lib.cubbs_hip_tune_load(b"pipeline kind=verify log2_items=13 chunk=4 group=8 super=1\n") == 1
    assert lib.cubbs_hip_tune_load(b"pipeline kind=verify log2_items=13 chunk=4 group=8 super=9\n") == 1
    assert lib.cubbs_hip_tune_load(b"pipeline kind=verify log2_items=13 chunk=4 group=8 super=x\n") == 1
    # signatures per product job of a Verify group: 0..4
    assert lib.cubbs_hip_tune_load(b"pipeline kind=verify log2_items=13 chunk=2 group=8 super=0 straus=1\n") == 0
    assert lib.cubbs_hip_tune_load(b"pipeline kind=verify log2_items=13 chunk=2 group=8 straus=4\n") == 0
    assert lib.cubbs_hip_tune_load(b"pipeline kind=verify log2_items=13 chunk=2 group=8 straus=5\n") == 1
    assert lib.cubbs_hip_tune_load(b"pipeline kind=verify log2_items=13 chunk=2 group=8 straus=y\n") == 1
    assert lib.cubbs_hip_tune_load(b"") == 0


def test_committed_pipeline_table_loads():
    # the table bench.py installs for the headline run (written by
    # bench.py --tune-pipeline on the GPU)
    path = os.path.join(ROOT, "cubbs-cred_amd", "pipeline_tuning.json")
    if not os.path.exists(path):
        pytest.skip("no pipeline tuning table committed")
    with open(path) as f:
        table = json.load(f)
    lib = cubbs_hip.lib()
    try:
        assert lib.cubbs_hip_tune_load("\n".join(table["table"]).encode()) == 0
    finally:
        lib.cubbs_hip_tune_load(b"")
    for c in table["classes"]:
        best = c["best"]
        assert best in c["candidates"]
        assert best["ms_per_step"] == min(x["ms_per_step"] for x in c["candidates"])


def test_prepare_flags_reject_bad_shapes_before_device():
    lib = cubbs_hip.lib()
    st = ctypes.c_int32(0)
    items = (cubbs_hip.VerifyItem * 1)()
    for flags in ((65 << 12), (127 << 12)):
        h = lib.cubbs_hip_batch_prepare_flags(2, cubbs_hip.BATCH_VERIFY, ctypes.cast(items, ctypes.c_void_p),
                                              ctypes.c_size_t(1), flags, ctypes.byref(st))
        assert not h and st.value == cubbs_hip.ERR_INVALID_INPUT
    # comb chunk on an MSM kind
    h = lib.cubbs_hip_batch_prepare_flags(2, cubbs_hip.BATCH_G1_MSM, ctypes.cast(items, ctypes.c_void_p),
                                          ctypes.c_size_t(1), 4 << 12, ctypes.byref(st))
    assert not h and st.value == cubbs_hip.ERR_INVALID_INPUT
    # supergroups: without aggregation, of one group, of more than 8 groups
    for flags in ((4 << 20), cubbs_hip.BATCH_AGGREGATE | (1 << 20), cubbs_hip.BATCH_AGGREGATE | (9 << 20)):
        h = lib.cubbs_hip_batch_prepare_flags(2, cubbs_hip.BATCH_VERIFY, ctypes.cast(items, ctypes.c_void_p),
                                              ctypes.c_size_t(1), flags, ctypes.byref(st))
        assert not h and st.value == cubbs_hip.ERR_INVALID_INPUT
