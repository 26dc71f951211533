"""Checks on the gfx950 assembly of the mixed-precision kernels (built here).

The FP16 lo-term split of pass 2 is inline asm (v_fma_mixlo/mixhi_f16), and
the compiler's hazard recognizer does not count inline asm as a VALU write: it
would insert no wait states before an MFMA that reads such a VGPR, although
the hardware needs two.  The kernel fences the split off with a scheduling
barrier and an s_nop; this test compiles the source to assembly and checks, for
every hi/lo instantiation of the pass-2 kernel, that no MFMA reads a VGPR
written by a v_fma_mix instruction fewer than three instructions earlier in
the same straight-line stretch (the single-term instantiations have no split).
"""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "da-gtm_amd", "csrc", "dg_mixed.hip")
HIPCC = "/opt/rocm/bin/hipcc"


def _regs(tok):
    m = re.match(r"v\[(\d+):(\d+)\]", tok)
    if m:
        return [f"v{r}" for r in range(int(m.group(1)), int(m.group(2)) + 1)]
    m = re.match(r"(v\d+)\b", tok)
    return [m.group(1)] if m else []


def _min_mix_to_mfma(lines):
    """Smallest instruction distance from a v_fma_mix write to an MFMA A/B read."""
    last, worst = {}, None
    for k, x in enumerate(lines):
        if x.startswith(("s_branch", "s_cbranch")) or x.endswith(":"):
            last = {}
            continue
        op, _, rest = x.partition(" ")
        toks = [t.strip() for t in rest.split(",")]
        if op.startswith("v_fma_mix"):
            for r in _regs(toks[0]):
                last[r] = k
        if "mfma" in op and len(toks) >= 3:
            for t in toks[1:3]:
                for r in _regs(t):
                    if r in last and (worst is None or k - last[r] < worst):
                        worst = k - last[r]
    return worst


@pytest.fixture(scope="module")
def asm(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    out = tmp_path_factory.mktemp("isa") / "dg_mixed.s"
    cmd = [HIPCC, "-O3", "-std=c++17", "--offload-arch=gfx950", "--cuda-device-only", "-S",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.dirname(SRC), SRC, "-o", str(out)]
    subprocess.run(cmd, check=True, capture_output=True, timeout=600)
    return out.read_text().split("\n")


def _kernels(lines, stem):
    out, cur = {}, None
    for l in lines:
        m = re.match(r"^(_Z\w*" + stem + r"\w*):", l)
        if m:
            cur = m.group(1)
            out[cur] = []
            continue
        if cur is None:
            continue
        x = l.strip()
        if x.startswith(".Lfunc_end"):
            cur = None
            continue
        if x and not x.startswith(";") and (not x.startswith(".") or re.match(r"^\.LBB\d+_\d+:", x)):
            out[cur].append(x)
    return out


def test_mix_split_is_fenced_from_mfma(asm):
    ks = {**_kernels(asm, "k_e2_mfma"), **_kernels(asm, "kb_e2_mfma")}
    assert len(ks) >= 64, "pass-2 instantiations not found"
    for name, body in ks.items():
        d = _min_mix_to_mfma(body)
        hilo = re.search(r"Lb[01]ELb1EE", name) is not None     # last template flag: R as hi/lo
        assert (d is not None) == hilo, name  # exactly the hi/lo kernels use the split
        assert d is None or d >= 3, (name, d)


def test_checker_sees_an_unfenced_read():
    """The checker itself: a v_fma_mix result read by the next MFMA is caught."""
    body = ["v_cvt_pk_f16_f32 v4, v1, v2",
            "v_fma_mixlo_f16 v3, v4, -1.0, v1 op_sel_hi:[1,0,0]",
            "v_mfma_f32_32x32x16_f16 a[0:15], v[20:23], v[0:3], a[0:15]"]
    assert _min_mix_to_mfma(body) == 1
    assert _min_mix_to_mfma(body[:2] + ["s_nop 1", "s_branch .LBB0_1", ".LBB0_1:"] + body[2:]) is None


def test_pass1_accumulator_in_vgprs(asm):
    """Pass 1 reads its accumulator with VALU FMAs: it must live in ordinary
    VGPRs (no AGPRs, no v_accvgpr_read), and every instantiation must still fit
    three waves per SIMD (at most 168 registers) -- except the screened
    single-instance kernel at D <= 192 (KS = 12), which is allowed two (at
    most 256)."""
    text = "\n".join(asm)
    names = sorted(set(re.findall(r"^(_Z\w*kb?_e1_mfma\w*):", text, re.M)))
    assert len(names) >= 30, "pass-1 instantiations not found"
    for name in names:
        agpr = re.search(r"\.set " + name + r"\.num_agpr, (\d+)", text)
        vgpr = re.search(r"\.set " + name + r"\.num_vgpr, (\d+)", text)
        assert agpr and vgpr, name
        assert int(agpr.group(1)) == 0, (name, agpr.group(1))
        cap = 256 if re.search(r"\dk_e1_mfmaILi12ELi1E", name) else 168
        assert int(vgpr.group(1)) <= cap, (name, vgpr.group(1))
    for stem in ("k_e1_mfma", "kb_e1_mfma"):
        for name, body in _kernels(asm, stem).items():
            assert not any(x.startswith("v_accvgpr_read") for x in body), name


def test_dma_uses_scalar_base(asm):
    """Every LDS-DMA of the E-pass streams takes its address as a scalar base
    plus a 32-bit lane offset (global_load_lds ... vN, s[a:b], or the SGPR pair
    vcc): no per-load
    64-bit VGPR address, which would cost a VALU add per load."""
    for stem in ("k_e1_mfma", "k_e2_mfma", "kb_e1_mfma", "kb_e2_mfma"):
        ks = _kernels(asm, stem)
        assert ks, stem
        for name, body in ks.items():
            loads = [x for x in body if x.startswith("global_load_lds")]
            assert loads, name
            bad = [x for x in loads
                   if not re.match(r"global_load_lds_dwordx4 v\d+, (s\[\d+:\d+\]|vcc)(,|\s|$)", x)]
            assert not bad, (name, bad[:2])


def test_pass2_loop_has_no_vector_scalar_loads(asm):
    """The per-block mask-threshold read of pass 2 is a scalar load: no
    single-dword vector load (global or flat) follows the first MFMA of a
    k_e2_mfma instantiation, so no such load's vmcnt wait (which would also
    wait for the LDS-DMAs) sits in the block loop.  In the persistent batched
    kernel the compiler cannot prove the thresholds unwritten across work
    items (the items' partial sums are stored through table pointers), so
    there the read stays a vector load: one per unrolled loop iteration and
    the prologue block's."""
    for stem, most in (("k_e2_mfma", 0), ("kb_e2_mfma", 4)):
        for name, body in _kernels(asm, stem).items():
            first = next(i for i, x in enumerate(body) if x.startswith("v_mfma"))
            n = sum(1 for x in body[first:]
                    if x.startswith(("global_load_dword ", "flat_load_dword ")))
            assert n <= most, (name, n)


def test_pass1_loop_waits_for_dma_only_at_the_barrier(asm):
    """Pass 1 reads everything it needs from LDS before it issues the tile's
    LDS-DMA, so every vmcnt wait in each instantiation belongs to a barrier
    (prologue and loop end): no LDS read sits between such a wait and the
    next s_barrier.  (Waits after the last barrier and LDS read -- the
    epilogue's reads of the centre-group maxima -- do not concern the loop.)"""
    for name, body in {**_kernels(asm, "k_e1_mfma"), **_kernels(asm, "kb_e1_mfma")}.items():
        for j, x in enumerate(body):
            if x.startswith("s_waitcnt") and "vmcnt" in x:
                nxt = next((y for y in body[j + 1:] if y.startswith(("s_barrier", "ds_read"))), None)
                assert nxt is None or nxt.startswith("s_barrier"), (name, body[j - 2:j + 3])


def test_dma_wait_checker_sees_an_early_wait():
    """The checker itself: a vmcnt wait followed by an LDS read before the
    barrier is what it must reject."""
    body = ["global_load_lds_dwordx4 v1, s[2:3]", "s_waitcnt vmcnt(0)", "ds_read_b128 v[4:7], v2", "s_barrier"]
    nxt = next((y for y in body[2:] if y.startswith(("s_barrier", "ds_read"))), None)
    assert nxt.startswith("ds_read")


def _e2_lds_reads_expected(KS, TT):
    """16-byte LDS reads of one pass-2 instantiation: the prologue block
    (KS TT data-row operands, 4 pB and 4 mask-threshold vectors), three
    unrolled loop iterations (the same plus the 2 DB TT transposed-row
    operands of the R T product) and the epilogue block's R T operands.
    With exact data and KS >= 4 the loop takes every hi-term MFMA before the
    lo-term ones and reads each data-row operand once per half (2 KS)."""
    DB = (16 * KS + 31) // 32
    loop_rows = 2 * KS if (TT == 1 and KS >= 4) else KS * TT
    return (KS * TT + 8) + 3 * (loop_rows + 2 * DB * TT + 8) + 2 * DB * TT


def test_pass2_reads_every_operand(asm):
    """Every pass-2 instantiation reads all of its R T A operands in each
    unrolled loop iteration: a schedule slot that never comes (for instance
    an operand load placed before the first cross-term MFMA of a phase with
    fewer MFMAs than the offset) leaves an operand register unread and the
    R T product reads stale data."""
    for name, body in {**_kernels(asm, "k_e2_mfma"), **_kernels(asm, "kb_e2_mfma")}.items():
        m = re.search(r"kb?_e2_mfmaILi(\d+)ELi(\d+)ELb([01])ELb([01])E", name)
        assert m, name
        KS, TT = int(m.group(1)), int(m.group(2))
        n = sum(1 for x in body if x.startswith("ds_read_b128"))
        assert n == _e2_lds_reads_expected(KS, TT), (name, n, _e2_lds_reads_expected(KS, TT))


def test_e_passes_do_not_spill(asm):
    """No E-pass instantiation, single or batched (work-queue loop), uses
    scratch memory: in the batched loop the bodies' lane-derived values are
    recomputed per work item rather than held across it."""
    text = "\n".join(asm)
    names = set(re.findall(r"^(_Z\w*kb?_e[12]_mfma\w*):", text, re.M))
    assert len(names) >= 90
    for name in names:
        ps = re.search(r"\.set " + re.escape(name) + r"\.private_seg_size, (\d+)", text)
        assert ps and int(ps.group(1)) == 0, (name, ps and ps.group(1))

