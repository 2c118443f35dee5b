"""The gfx950 code of every kernel file, compiled to AMD assembly here, holds
none of the instructions that would round differently from the arithmetic
as written: mixed-precision FMAs with an fp16 result (a multiply-add fused
into the fp16 conversion, one rounding where the source asks for an fp32
rounding first), the non-fused mixed mad, and packed FP32 math.
v_fma_mix_f32 stays allowed: it is an fp32 fma whose fp16 inputs are
converted exactly, the same result as the explicit fmaf it comes from.

Every kernel also ends in s_endpgm: undefined behaviour in a kernel (a shift
by the operand's width or more, say) lets the compiler drop the whole body,
and a body without an end runs into whatever code follows it."""
import concurrent.futures
import os
import re
import shutil
import subprocess
import tempfile

import pytest

from restir_amd import build as rb

FORBIDDEN = re.compile(r"\bv_(fma_mixlo_f16|fma_mixhi_f16|mad_mix\w*|pk_fma_f32|pk_mul_f32|pk_add_f32)\b")


def device_asm(src, out):
    cmd = [rb.HIPCC] + rb.HIP_FLAGS + ["-S", "--offload-device-only", os.path.join(rb.CSRC, src), "-o", out]
    subprocess.run(cmd, check=True, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    return open(out).read()


@pytest.mark.skipif(shutil.which(rb.HIPCC) is None and not os.path.exists(rb.HIPCC), reason="no hipcc")
def test_no_fused_or_packed_fp32_in_kernels():
    tmp = tempfile.mkdtemp()
    try:
        with concurrent.futures.ThreadPoolExecutor(4) as pool:
            outs = list(pool.map(lambda s: (s, device_asm(s, os.path.join(tmp, s + ".s"))), rb.HIP_SOURCES))
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    for src, text in outs:
        code = [l for l in text.splitlines() if not l.lstrip().startswith((";", "."))]
        bad = [l.strip() for l in code if FORBIDDEN.search(l)]
        assert not bad, "%s: %s" % (src, bad[:5])
        kernels = re.findall(r"^\s*\.amdhsa_kernel (\S+)", text, re.M)
        for k in kernels:
            start = text.index("\n" + k + ":")
            end = text.index(".Lfunc_end", start)
            assert "s_endpgm" in text[start:end], "%s: kernel %s has no s_endpgm" % (src, k)
