"""The library's host code on the CPU, through a stand-in HIP runtime.

libmtcheck_fakehip.so is the library's own objects linked against
tests/fakehip/fake_hip.cpp: device memory is host memory, kernels never run, and
every launch is recorded.  The pointer arguments of each batched launch are
checked against the live allocations (each required pointer non-null, the bytes
a kernel may touch inside one allocation).  These tests run whole checkpoint
steps -- tree and basic members, baselines, graph replays, batches -- and
require that no launch breaks that rule, and they check the host-side error
paths (limits, refused steps, failed steps).  CPU only.
"""
import ctypes

import pytest

from tests.fakehip import build_fakehip

TREE, BASIC = 6, 2
OK, INVALID_ARG, OOM, TABLE_FULL, STATE = 0, 1, 3, 4, 7


class Dim3(ctypes.Structure):
    _fields_ = [("x", ctypes.c_uint32), ("y", ctypes.c_uint32), ("z", ctypes.c_uint32)]


class Header(ctypes.Structure):
    _fields_ = [("ref_id", ctypes.c_uint32), ("chkpt_id", ctypes.c_uint32), ("datalen", ctypes.c_uint64),
                ("chunk_size", ctypes.c_uint32), ("num_first_ocur", ctypes.c_uint32),
                ("num_prior_chkpts", ctypes.c_uint32), ("num_shift_dupl", ctypes.c_uint32)]


@pytest.fixture(scope="module")
def fake():
    path = build_fakehip.build()
    L = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)
    L.fakehip_violation.restype = ctypes.c_char_p
    L.fakehip_launch_name.restype = ctypes.c_char_p
    L.fakehip_bytes_alive.restype = ctypes.c_uint64
    L.mtc_last_error.restype = ctypes.c_char_p
    L.mtc_batch_last_error.restype = ctypes.c_char_p
    return L


def violations(L):
    return [L.fakehip_violation(i).decode() for i in range(L.fakehip_violation_count())]


def launches(L):
    return [L.fakehip_launch_name(i).decode() for i in range(L.fakehip_launch_count())]


def dev_alloc(L, size):
    p = ctypes.c_void_p()
    assert L.hipMalloc(ctypes.byref(p), ctypes.c_size_t(size)) == 0
    return p


def create(L, C, approach=TREE, hk=0):
    h = ctypes.c_void_p()
    assert L.mtc_create(ctypes.byref(h), C, approach, hk, 0) == OK
    return h


def checkpoint(L, h, data, size, baseline=False):
    hdr = Header()
    sz = ctypes.c_uint64()
    rc = L.mtc_checkpoint(h, data, ctypes.c_uint64(size), int(baseline), 1, None, ctypes.byref(hdr), ctypes.byref(sz))
    return rc, hdr


def test_checker_flags_the_unallocated_table_reset(fake):
    """Negative control: the launch a basic baseline made before its table reset
    was skipped -- k_fill_u64 over 4 words of a null table -- is reported."""
    L = fake
    L.fakehip_reset()
    stub = getattr(L, "_ZN3mtc25__device_stub__k_fill_u64EPmmm")  # the host stub of k_fill_u64
    L.__hipPushCallConfiguration(Dim3(1, 1, 1), Dim3(256, 1, 1), ctypes.c_size_t(0), None)
    stub(None, ctypes.c_uint64(4), ctypes.c_uint64(~0 & (2**64 - 1)))
    v = violations(L)
    assert len(v) == 1 and "k_fill_u64" in v[0] and "null" in v[0], v
    small = dev_alloc(L, 64)
    L.fakehip_reset()
    L.__hipPushCallConfiguration(Dim3(1, 1, 1), Dim3(256, 1, 1), ctypes.c_size_t(0), None)
    stub(small, ctypes.c_uint64(9), ctypes.c_uint64(0))  # 72 bytes into a 64-byte allocation
    assert any("extends past" in x for x in violations(L))
    L.fakehip_reset()


@pytest.mark.parametrize("C,size", [(128, 1000 * 128 + 5), (64, 64 * 64), (16, 16 * 3 + 1), (100, 100 * 777)])
def test_basic_steps_launch_in_bounds(fake, C, size):
    L = fake
    L.fakehip_reset()
    h = create(L, C, BASIC)
    data = dev_alloc(L, size)
    for k in range(5):
        rc, hdr = checkpoint(L, h, data, size, baseline=(k == 0 or k == 3))
        assert rc == OK, L.mtc_last_error(h)
        assert (hdr.ref_id, hdr.chkpt_id) == ((3 if k >= 3 else 0), k)
    names = launches(L)
    assert any("basic_diff" in n for n in names)
    assert not any("k_fill_u64" in n for n in names), "a basic handle keeps no table to reset"
    assert violations(L) == []
    L.mtc_destroy(h)


@pytest.mark.parametrize("C,size,hk,tuning", [
    (128, 1000 * 128 + 5, 0, None),
    (128, (1 << 16) * 128 + 77, 1, None),
    (1, 3, 0, None),
    (128, 128, 0, None),
    (64, 64 * 5000, 0, (4, 0.25, 3)),
    (100, 100 * 777, 1, (2, 0.125, 5)),
    (128, 128 * 70000 + 9, 0, (1, 0.25, 8, 64, 3)),
    (128, 128 * 40000, 1, (1, 0.125, 7, 1, 1)),
    (4096, 4096 * 300, 0, None),
])
def test_tree_steps_launch_in_bounds(fake, C, size, hk, tuning):
    L = fake
    L.fakehip_reset()
    h = create(L, C, TREE, hk)
    if tuning:
        assert L.mtc_set_tuning(h, tuning[0], ctypes.c_float(tuning[1]), tuning[2]) == OK
        if len(tuning) > 3:
            assert L.mtc_set_sweep_tiles(h, tuning[3], tuning[4]) == OK
    assert L.mtc_set_sweep_tiles(h, 65, 0) == INVALID_ARG and L.mtc_set_sweep_tiles(h, 0, 65) == INVALID_ARG
    data = dev_alloc(L, size)
    for k in range(4):
        if k == 2:
            L.mtc_set_debug(h, 1)
        rc, hdr = checkpoint(L, h, data, size, baseline=(k == 0))
        assert rc == OK, L.mtc_last_error(h)
    names = launches(L)
    assert any("assemble" in n for n in names)
    if size // C >= 300:  # two sweep stages or more: the last small ones run as one block
        assert any("9SweepTail" in n for n in names)
    assert violations(L) == []
    L.mtc_destroy(h)


def test_leaf_labels_written_for_listed_chunks_once_reset(fake, monkeypatch):
    """The leaf kernel labels every chunk until a completed step has reset the
    labels of the chunks it listed (k_assemble does, unless the debug view needs
    them); after that it labels only the chunks it lists.  A debug step, a failed
    step and a basic handle keep every chunk labelled."""
    L = fake
    L.fakehip_reset()
    h = create(L, 128)
    size = 128 * 3000
    data = dev_alloc(L, size)

    def step(baseline=False, expect=OK):
        n0 = 0
        while L.fakehip_leaf_all_labels(n0) >= 0:
            n0 += 1
        r0 = 0
        while L.fakehip_label_resets(r0) >= 0:
            r0 += 1
        rc, _ = checkpoint(L, h, data, size, baseline=baseline)
        assert rc == expect, L.mtc_last_error(h)
        return L.fakehip_leaf_all_labels(n0), L.fakehip_label_resets(r0)

    assert step(baseline=True) == (1, 1)   # fresh labels: all written, listed ones reset
    assert step() == (0, 1)
    assert step() == (0, 1)
    L.mtc_set_debug(h, 1)
    assert step() == (1, 0)                # the debug view copies the labels after the step
    L.mtc_set_debug(h, 0)
    assert step() == (1, 1)                # the debug step left them unreset
    assert step() == (0, 1)
    monkeypatch.setenv("MTC_TEST_FAIL_MEMBER", "0")
    assert step(expect=STATE) == (0, 1)
    monkeypatch.delenv("MTC_TEST_FAIL_MEMBER")
    assert step() == (1, 1)                # a failed step: a baseline, every label written
    assert violations(L) == []
    L.mtc_destroy(h)
    b = create(L, 128, BASIC)
    rc, _ = checkpoint(L, b, data, size, baseline=True)
    assert rc == OK
    rc, _ = checkpoint(L, b, data, size)
    assert rc == OK
    n = 0
    while L.fakehip_leaf_all_labels(n) >= 0:
        n += 1
    assert L.fakehip_leaf_all_labels(n - 1) == 1 and L.fakehip_leaf_all_labels(n - 2) == 1
    L.mtc_destroy(b)


def test_listed_digests_stored_or_hashed_again(fake):
    """An incremental step of a buffer of at least 2^21 chunks of 128 bytes after one
    that changed few chunks stores no digest in the leaf kernel and hashes the
    listed chunks in k_compact instead (Murmur3, one tile per leaf block, aligned
    data); a baseline, a smaller buffer, other chunk sizes, MD5 and several tiles
    per leaf block store them.  (No kernel runs here, so the previous step's changed
    count reads zero.)"""
    L = fake
    big = 1 << 21

    def flags(h, size, baseline):
        L.fakehip_reset()
        data = dev_alloc(L, size)
        rc, _ = checkpoint(L, h, data, size, baseline=baseline)
        assert rc == OK, L.mtc_last_error(h)
        assert violations(L) == []
        leaf, comp = [], []
        while L.fakehip_leaf_hash_later(len(leaf)) >= 0:
            leaf.append(L.fakehip_leaf_hash_later(len(leaf)))
        while L.fakehip_compact_hash(len(comp)) >= 0:
            comp.append(L.fakehip_compact_hash(len(comp)))
        return leaf, comp

    h = create(L, 128)
    assert flags(h, 128 * big, True) == ([0], [0])
    assert flags(h, 128 * big, False) == ([1], [1])
    assert flags(h, 128 * 3000, True) == ([0], [0])
    assert flags(h, 128 * 3000, False) == ([0], [0])
    assert L.mtc_set_tuning(h, 4, ctypes.c_float(0.125), 8) == OK
    assert flags(h, 128 * big, True) == ([0], [0])
    assert flags(h, 128 * big, False) == ([0], [0])
    L.mtc_destroy(h)
    for C, hk in [(128, 1), (100, 0), (64, 0), (4096, 0)]:
        h = create(L, C, hk=hk)
        assert flags(h, C * big + 7, True) == ([0], [0])
        assert flags(h, C * big + 7, False) == ([0], [0])
        L.mtc_destroy(h)
    h = create(L, 128)
    assert flags(h, 128 * big + 5, True) == ([0], [0])
    assert flags(h, 128 * big + 5, False) == ([1], [1])  # a short last chunk: hashed byte by byte
    L.mtc_destroy(h)


def test_batch_of_tree_and_basic_members_with_graph_replay(fake):
    L = fake
    L.fakehip_reset()
    b = ctypes.c_void_p()
    assert L.mtc_batch_create(ctypes.byref(b), 0, 1) == OK
    specs = [(128, TREE, 0, 128 * 5000 + 3), (64, BASIC, 0, 64 * 999), (512, TREE, 1, 512 * 1234),
             (128, BASIC, 0, 128 * 77 + 9), (16, TREE, 0, 16 * 70000)]
    hs = [create(L, C, a, hk) for C, a, hk, _ in specs]
    datas = [dev_alloc(L, s) for *_, s in specs]
    B = len(specs)
    H = (ctypes.c_void_p * B)(*[h.value for h in hs])
    D = (ctypes.c_void_p * B)(*[d.value for d in datas])
    N = (ctypes.c_uint64 * B)(*[s for *_, s in specs])
    hdrs = (Header * B)()
    sizes = (ctypes.c_uint64 * B)()
    for k in range(4):
        base = (ctypes.c_int * B)(*([1] * B if k == 0 else [0] * B))
        rc = L.mtc_checkpoint_batch(b, H, D, N, base, B, 1, None, hdrs, sizes)
        assert rc == OK, L.mtc_batch_last_error(b)
        assert [x.chkpt_id for x in hdrs] == [k] * B
    assert any("basic_diff" in n for n in launches(L))
    assert violations(L) == []
    for h in hs:
        L.mtc_destroy(h)
    L.mtc_batch_destroy(b)


def test_batch_limits(fake):
    L = fake
    b = ctypes.c_void_p()
    assert L.mtc_batch_create(ctypes.byref(b), 0, 0) == OK
    h = create(L, 128)
    data = dev_alloc(L, 4096)
    count = 65537
    H = (ctypes.c_void_p * count)(*([h.value] * count))
    D = (ctypes.c_void_p * count)(*([data.value] * count))
    N = (ctypes.c_uint64 * count)(*([4096] * count))
    assert L.mtc_checkpoint_batch(b, H, D, N, None, count, 1, None, None, None) == INVALID_ARG
    assert b"65536" in L.mtc_batch_last_error(b)
    # the same handle twice in one batch
    assert L.mtc_checkpoint_batch(b, H, D, N, None, 2, 1, None, None, None) == INVALID_ARG
    assert b"twice" in L.mtc_batch_last_error(b)
    L.mtc_destroy(h)
    L.mtc_batch_destroy(b)


def test_table_sizing_limit_refuses_before_allocating(fake):
    """2^28 + 1 one-byte chunks need 2^32 table slots at load factor 0.125: slot
    indices are 32-bit, so the checkpoint is refused before anything is allocated."""
    L = fake
    h = create(L, 1)
    data = dev_alloc(L, 1)  # never read: the step is refused first
    before = L.fakehip_bytes_alive()
    rc, _ = checkpoint(L, h, data, (1 << 28) + 1, baseline=True)
    assert rc == TABLE_FULL, L.mtc_last_error(h)
    assert L.fakehip_bytes_alive() == before
    # 2^28 chunks still fit (2^31 slots)
    h2 = create(L, 1)
    L.fakehip_reset()
    rc, _ = checkpoint(L, h2, data, 1 << 28, baseline=True)
    assert rc == OK, L.mtc_last_error(h2)
    assert violations(L) == [] or all("data extends past" in v for v in violations(L))
    L.mtc_destroy(h2)
    L.mtc_destroy(h)


def test_refused_step_leaves_handles_unchanged(fake):
    """A batch refused at validation (one member's length changed) changes no
    handle: the other member's next checkpoint is still incremental."""
    L = fake
    b = ctypes.c_void_p()
    assert L.mtc_batch_create(ctypes.byref(b), 0, 0) == OK
    a, c = create(L, 128), create(L, 128)
    da, dc = dev_alloc(L, 128 * 100), dev_alloc(L, 128 * 200)
    assert checkpoint(L, a, da, 128 * 100, True)[0] == OK
    assert checkpoint(L, c, dc, 128 * 200, True)[0] == OK
    H = (ctypes.c_void_p * 2)(a.value, c.value)
    D = (ctypes.c_void_p * 2)(da.value, dc.value)
    N = (ctypes.c_uint64 * 2)(128 * 100, 128 * 150)  # c's length differs from its previous checkpoint
    assert L.mtc_checkpoint_batch(b, H, D, N, None, 2, 1, None, None, None) == STATE
    rc, hdr = checkpoint(L, a, da, 128 * 100)
    assert rc == OK and (hdr.ref_id, hdr.chkpt_id) == (0, 1)
    L.mtc_destroy(a)
    L.mtc_destroy(c)
    L.mtc_batch_destroy(b)


def test_failed_step_drops_every_members_state(fake, monkeypatch):
    """A member whose kernels report an error fails the batch; every member of that
    step then starts a new baseline (the step already touched all of them)."""
    L = fake
    b = ctypes.c_void_p()
    assert L.mtc_batch_create(ctypes.byref(b), 0, 1) == OK
    hs = [create(L, 128) for _ in range(3)]
    ds = [dev_alloc(L, 128 * 300) for _ in range(3)]
    H = (ctypes.c_void_p * 3)(*[h.value for h in hs])
    D = (ctypes.c_void_p * 3)(*[d.value for d in ds])
    N = (ctypes.c_uint64 * 3)(*([128 * 300] * 3))
    hdrs = (Header * 3)()
    assert L.mtc_checkpoint_batch(b, H, D, N, None, 3, 1, None, hdrs, None) == OK
    monkeypatch.setenv("MTC_TEST_FAIL_MEMBER", "1")
    assert L.mtc_checkpoint_batch(b, H, D, N, None, 3, 1, None, hdrs, None) == STATE
    monkeypatch.delenv("MTC_TEST_FAIL_MEMBER")
    assert L.mtc_checkpoint_batch(b, H, D, N, None, 3, 1, None, hdrs, None) == OK
    # ids did not advance on the failed step; each member restarted from a baseline
    assert [(x.ref_id, x.chkpt_id) for x in hdrs] == [(1, 1)] * 3
    for h in hs:
        L.mtc_destroy(h)
    L.mtc_batch_destroy(b)


class LeafArgs(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("len", ctypes.c_uint64), ("n", ctypes.c_uint32), ("C", ctypes.c_uint32),
                ("leaf_dg", ctypes.c_void_p), ("labels", ctypes.c_void_p), ("baseline", ctypes.c_int),
                ("list", ctypes.c_void_p), ("blk_cnt", ctypes.c_void_p), ("tiles_per_block", ctypes.c_uint64),
                ("all_labels", ctypes.c_int), ("list_dg", ctypes.c_void_p), ("stamp", ctypes.c_void_p),
                ("hash_later", ctypes.c_int)]


class FillSeg(ctypes.Structure):
    _fields_ = [("p", ctypes.c_void_p), ("count", ctypes.c_uint64), ("v", ctypes.c_uint32)]


class Prologue(ctypes.Structure):
    _fields_ = [("src", ctypes.c_void_p), ("dst", ctypes.c_void_p), ("n16", ctypes.c_uint64),
                ("fill", FillSeg * 12), ("nfill", ctypes.c_uint32)]


def test_checker_flags_leaf_counts_not_known_zero(fake):
    """Negative control: a direct leaf launch whose block counts were neither filled
    nor cleared by a k_compact is reported, and so is one that fills them itself."""
    L = fake
    L.fakehip_reset()
    n = 1000
    a = LeafArgs(dev_alloc(L, 128 * n), 128 * n, n, 128, dev_alloc(L, 16 * n), dev_alloc(L, n), 0,
                 dev_alloc(L, 4 * n), dev_alloc(L, 4 * 4), 1, 0, dev_alloc(L, 16 * n), None, 0)
    stub = getattr(L, "_ZN3mtc29__device_stub__k_leaf1_directILi0ELi128EEEvNS_8LeafArgsENS_8PrologueE")
    L.__hipPushCallConfiguration(Dim3(4, 1, 1), Dim3(256, 1, 1), ctypes.c_size_t(0), None)
    stub(a, Prologue())
    assert any("not known to be zero" in v for v in violations(L)), violations(L)
    L.fakehip_reset()
    pro = Prologue()
    pro.fill[0] = FillSeg(a.blk_cnt, 4, 0)  # a fill inside the same launch does not count
    pro.nfill = 1
    L.__hipPushCallConfiguration(Dim3(4, 1, 1), Dim3(256, 1, 1), ctypes.c_size_t(0), None)
    stub(a, pro)
    v = violations(L)
    assert any("not known to be zero" in x for x in v) and any("fills the block counts" in x for x in v), v
    L.fakehip_reset()


def test_direct_leaf_launch_follows_cleared_counts(fake, monkeypatch):
    """A one-member step launches its leaf kernel by value (no upload copy, no fill
    kernel) only when the last completed step's k_compact cleared the block counts;
    after a failed step it fills them again first."""
    L = fake
    L.fakehip_reset()
    h = create(L, 128)
    size = 128 * 5000
    data = dev_alloc(L, size)
    kinds = []
    for k in range(6):
        L.fakehip_reset()
        if k == 3:
            monkeypatch.setenv("MTC_TEST_FAIL_MEMBER", "0")
        rc, _ = checkpoint(L, h, data, size, baseline=(k == 0))
        monkeypatch.delenv("MTC_TEST_FAIL_MEMBER", raising=False)
        assert rc == (STATE if k == 3 else OK), L.mtc_last_error(h)
        names = launches(L)
        kinds.append("direct" if any("k_leaf1_direct" in x for x in names) else
                     "filled" if any("k_fill" in x for x in names) else "?")
        assert violations(L) == [], violations(L)
    # k == 4 runs after the failed step: a new baseline, counts filled again
    assert kinds == ["filled", "direct", "direct", "direct", "filled", "direct"], kinds
    L.mtc_destroy(h)


def test_server_admission_and_submit_limits(fake):
    L = fake
    s = ctypes.c_void_p()
    assert L.mtc_server_create(ctypes.byref(s), 0, ctypes.c_uint64(64 << 20), 1, 4) == OK
    assert L.mtc_server_create(ctypes.byref(ctypes.c_void_p()), 0, ctypes.c_uint64(1 << 30), 1, 65537) == INVALID_ARG
    cid = ctypes.c_uint32()
    # 1 GiB of 128-byte chunks needs about 4 GiB of state: it could never be admitted
    assert L.mtc_server_add_client(s, 128, 0, ctypes.c_uint64(1 << 30), ctypes.byref(cid)) == OOM
    assert L.mtc_server_add_client(s, 128, 0, ctypes.c_uint64(1 << 20), ctypes.byref(cid)) == OK
    data = dev_alloc(L, 2 << 20)
    t = ctypes.c_uint64()
    assert L.mtc_server_submit(s, cid, data, ctypes.c_uint64(2 << 20), 1, None, ctypes.byref(t)) == INVALID_ARG
    assert L.mtc_server_submit(s, cid, data, ctypes.c_uint64(1 << 20), 1, None, ctypes.byref(t)) == OK
    hdr = Header()
    assert L.mtc_server_result(s, t, 1, ctypes.byref(hdr), None, None, None, None) == OK
    assert L.mtc_server_destroy(s) == OK
