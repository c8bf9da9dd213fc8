"""The .dat V6 store through the C ABI (host code only, no GPU): byte
compatibility with the format (checked against the oracle's byte image, which
test_oracle.py pins to DATA_FORMAT.md's hex examples), reader fast and scan
paths, resume, the partial-record fix, the reference's error messages,
remaining seeds and percentiles."""
import os
import struct

import numpy as np
import pytest


def recs(seeds, nf, seed0=0):
    rng = np.random.default_rng(seed0)
    return [(int(s), rng.random(nf).tolist()) for s in seeds]


def write(jne, path, model, dim, steps, records, finish=True, quiet=True):
    w = jne.DatWriter(path, model, dim, steps, quiet=quiet)
    if records:
        w.append([r[0] for r in records], np.array([r[1] for r in records]))
    if finish:
        w.finish()
    else:
        w.abandon()


def test_writer_bytes_match_format(jne, O, tmp_path):
    p = tmp_path / "a.dat"
    rs = recs([1, 2, 127, 128, 300, 16384, 2**21, 2**28, 0xFFFFFFFF], 3)
    write(jne, p, 4, 2, 10000, rs)
    assert p.read_bytes() == O.dat_bytes(4, 2, 10000, rs)
    # empty file: header + trailer with count 0 and nF 0
    q = tmp_path / "e.dat"
    write(jne, q, 0, 1, 10, [])
    assert q.read_bytes() == O.dat_bytes(0, 1, 10, [])


def test_reader_fast_and_scan_paths(jne, O, tmp_path):
    rs = recs(range(1, 40), 5, 1)
    p = tmp_path / "f.dat"
    p.write_bytes(O.dat_bytes(2, 5, 777, rs))
    seeds, eigs, m, d, st = jne.read_dat(p)
    assert (m, d, st) == (2, 5, 777)
    assert seeds.tolist() == [r[0] for r in rs]
    assert np.array_equal(eigs, np.array([r[1] for r in rs]))
    info = jne.dat_info(p)
    assert info["has_trailer"] and info["records"] == 39 and info["nf"] == 5
    # no trailer: whole records are scanned, a cut-off record is ignored
    raw = O.dat_bytes(2, 5, 777, rs, finish=False)
    p.write_bytes(raw[:-13])
    seeds, eigs, *_ = jne.read_dat(p)
    assert seeds.tolist() == [r[0] for r in rs[:-1]]
    assert not jne.dat_info(p)["has_trailer"]
    # smaller than header + trailer: no records
    p.write_bytes(O.dat_bytes(1, 3, 5, [], finish=False))
    assert jne.dat_info(p)["records"] == 0


def test_reader_errors(jne, O, tmp_path):
    p = tmp_path / "x.dat"
    p.write_bytes(b"EIGENVALS_V5" + bytes(30))
    with pytest.raises(jne.JneError) as e:
        jne.read_dat(p)
    assert "File format error: magic header mismatch" in str(e.value) and e.value.code == jne.JNE_ERR_FORMAT
    with pytest.raises(jne.JneError) as e:
        jne.read_dat(tmp_path / "missing.dat")
    assert e.value.code == jne.JNE_ERR_IO and "No such file or directory" in str(e.value)
    # trailer promises more records than the file holds (reader.rs fast path fails)
    rs = recs([1, 2, 3], 2)
    raw = bytearray(O.dat_bytes(0, 2, 10, rs))
    raw[-9:-1] = struct.pack("<Q", 5)
    p.write_bytes(bytes(raw))
    with pytest.raises(jne.JneError):
        jne.read_dat(p)
    # a record whose count differs from the trailer's
    raw = bytearray(O.dat_bytes(0, 2, 10, rs))
    raw[-1] = 3
    p.write_bytes(bytes(raw))
    with pytest.raises(jne.JneError) as e:
        jne.read_dat(p)
    assert "Eigenvalue count mismatch: expected 3, actual 2" in str(e.value)


def test_resume_appends_after_trailer(jne, O, tmp_path, capfd):
    p = tmp_path / "r.dat"
    a, b = recs([1, 2, 3], 4, 2), recs([7, 5], 4, 3)
    write(jne, p, 3, 3, 50, a)
    w = jne.DatWriter(p, 3, 3, 50, quiet=False)
    assert w.count == 3
    w.append([r[0] for r in b], np.array([r[1] for r in b]))
    w.finish()
    out = capfd.readouterr().out
    assert "Detected existing file with 3 data records" in out
    assert "Removed EOF marker to enable append mode" in out
    assert "SUCCESS: append write completed, wrote 5 data records for model 3, dim 3, steps 50" in out
    assert p.read_bytes() == O.dat_bytes(3, 3, 50, a + b)


def test_resume_truncates_partial_record(jne, O, tmp_path, capfd):
    """A crash mid-record leaves a partial record; resuming drops it, so
    the appended records stay readable (the reference appends after it)."""
    p = tmp_path / "c.dat"
    a, b = recs(range(1, 11), 3, 4), recs([11, 12], 3, 5)
    raw = O.dat_bytes(1, 2, 100, a, finish=False)
    p.write_bytes(raw[:-7])  # record 10 cut in the middle
    w = jne.DatWriter(p, 1, 2, 100, quiet=False)
    assert w.count == 9
    w.append([r[0] for r in b], np.array([r[1] for r in b]))
    w.finish()
    assert "Truncated 19 bytes after the last whole record" in capfd.readouterr().out  # 26-byte record cut by 7
    assert p.read_bytes() == O.dat_bytes(1, 2, 100, a[:9] + b)
    # an abandoned writer leaves a trailer-less file that resumes the same way
    w = jne.DatWriter(p, 1, 2, 100)
    w.append([13], np.array([[0.5, 0.25, 0.125]]))
    w.abandon()
    seeds, *_ = jne.read_dat(p)
    assert seeds.tolist() == [r[0] for r in a[:9] + b] + [13]


def test_open_errors_and_recreate(jne, O, tmp_path, capfd):
    p = tmp_path / "m.dat"
    write(jne, p, 2, 4, 1000, recs([1], 4))
    for args, msg in [((1, 4, 1000), "Model mismatch: file has model 2, expected 1"),
                      ((2, 5, 1000), "Dimension mismatch: file has dim 4, expected 5"),
                      ((2, 4, 999), "Steps mismatch: file has steps 1000, expected 999")]:
        with pytest.raises(jne.JneError) as e:
            jne.DatWriter(p, *args)
        assert msg in str(e.value) and e.value.code == jne.JNE_ERR_FORMAT
    # another format version is replaced by a fresh file
    p.write_bytes(b"EIGENVALS_V5" + bytes(40))
    w = jne.DatWriter(p, 2, 4, 1000, quiet=False)
    assert w.count == 0
    w.finish()
    assert "WARNING: Incompatible file format detected, recreating file..." in capfd.readouterr().out
    assert p.read_bytes() == O.dat_bytes(2, 4, 1000, [])
    # new file with an expected size prints the reference's notice
    jne.DatWriter(tmp_path / "n.dat", 0, 1, 10, expected_records=5, quiet=False).finish()
    assert "File pre-allocation disabled to ensure resume functionality works correctly" in capfd.readouterr().out


def test_append_validation(jne, tmp_path):
    w = jne.DatWriter(tmp_path / "v.dat", 4, 3, 10)
    w.append([1], np.zeros((1, 3)))
    with pytest.raises(jne.JneError) as e:
        w.append([2], np.zeros((1, 4)))
    assert "Eigenvalue count mismatch: expected 3, actual 4 (model 4, dim 3, steps 10)" in str(e.value)
    with pytest.raises(jne.JneError) as e:
        w.append([3], np.zeros((1, 256)))
    assert "Too many eigenvalues: 256 exceeds maximum of 255" in str(e.value)
    w.finish()
    # a record without eigenvalues would not read back (the reader rejects a zero count)
    w = jne.DatWriter(tmp_path / "z.dat", 4, 3, 10)
    with pytest.raises(jne.JneError) as e:
        w.append([1], np.zeros((1, 0)))
    assert "Invalid eigenvalue count: cannot be zero" in str(e.value)
    w.abandon()


def test_255_eigenvalues_round_trip(jne, O, tmp_path):
    """append_writer_test.rs: 255 eigenvalues per record is the largest count."""
    path = tmp_path / "b.dat"
    e = np.arange(2 * 255, dtype=np.float64).reshape(2, 255) / 7.0
    w = jne.DatWriter(path, 2, 255, 11)
    w.append([1, 300], e)
    w.finish()
    seeds, got = jne.read_dat(path)[:2]
    assert seeds.tolist() == [1, 300] and np.array_equal(got, e)
    assert path.read_bytes() == O.dat_bytes(2, 255, 11, [(1, e[0].tolist()), (300, e[1].tolist())])


def test_dim_stored_as_u8(jne, O, tmp_path):
    """dim is a u8 in the header: 300 wraps to 44 (thread_manager.rs:52)."""
    p = tmp_path / "w.dat"
    write(jne, p, 0, 300, 5, [])
    assert p.read_bytes()[13] == 300 & 0xFF


def test_remaining_seeds(jne, O, tmp_path):
    p = tmp_path / "s.dat"
    rem, done = jne.remaining_seeds(p, 0, 2, 10, 6)
    assert rem.tolist() == [1, 2, 3, 4, 5, 6] and done == 0
    write(jne, p, 0, 2, 10, recs([5, 1, 2, 9], 2))
    rem, done = jne.remaining_seeds(p, 0, 2, 10, 6)
    assert rem.tolist() == [3, 4, 6] and done == 4
    with pytest.raises(jne.JneError) as e:
        jne.remaining_seeds(p, 1, 2, 10, 6)
    assert "Model mismatch" in str(e.value)
    # an unreadable file counts as empty (progress.rs:52)
    p.write_bytes(b"garbage")
    assert jne.remaining_seeds(p, 0, 2, 10, 3)[0].tolist() == [1, 2, 3]


def test_percentiles(jne, O):
    """rank p*(n-1) interpolation (simulation_analyzers.rs:4-18), against the oracle and numpy."""
    v = [5.0, 1.0, 4.0, 2.0, 3.0]
    assert jne.percentiles(v, [0.0, 0.5, 1.0, 0.25, 0.9]) == [1.0, 3.0, 5.0, 2.0, pytest.approx(4.6)]
    rng = np.random.default_rng(3)
    x = rng.standard_normal(1001)
    ps = [0.5, 0.75, 0.8, 0.85, 0.9, 0.95, 0.975, 0.99]
    got = jne.percentiles(x, ps)
    assert got == pytest.approx(O.percentiles(x, ps), rel=0, abs=0)
    assert got == pytest.approx(np.percentile(x, [100 * p for p in ps], method="linear"), rel=1e-15)
    assert np.isnan(jne.percentiles([], [0.5])[0])
    with pytest.raises(jne.JneError):
        jne.percentiles([1.0, float("nan")], [0.5])


def test_statistics_from_file(jne, O, tmp_path):
    import ctypes

    p = tmp_path / "t.dat"
    rs = recs([3, 1, 2, 8], 3, 9)
    write(jne, p, 2, 3, 10, rs)
    L = jne.load_library()
    for stat, f in [(jne.STAT_TRACE, sum), (jne.STAT_MAXEIG, max)]:
        out = np.empty(4)
        n = ctypes.c_uint64()
        assert L.jne_dat_statistics(os.fsencode(p), 3, stat, out.ctypes.data, 4, ctypes.byref(n)) == 0
        assert n.value == 3  # seed 8 > runs 3 is left out
        assert out[:3].tolist() == [f(r[1]) for r in rs if r[0] <= 3]


def test_truncated_and_corrupted_files(jne, tmp_path):
    """Every truncation of a small file and a few hundred random byte
    corruptions: reading either succeeds or raises the library's error, and a
    writer opened on the damaged file (resume) leaves a file that reads back."""
    base = tmp_path / "base.dat"
    seeds = np.array([1 + 997 * i for i in range(30)] + [0xFFFFFFFF], dtype=np.uint32)
    eigs = np.arange(len(seeds) * 3, dtype=np.float64).reshape(-1, 3) / 3.0
    w = jne.DatWriter(base, 4, 3, 77)
    w.append(seeds, eigs)
    w.finish()
    good = base.read_bytes()
    cur = tmp_path / "cur.dat"
    rng = np.random.default_rng(3)
    cases = [good[:n] for n in range(len(good) + 1)]
    for _ in range(300):
        b = bytearray(good)
        for _ in range(int(rng.integers(1, 5))):
            b[int(rng.integers(0, len(b)))] = int(rng.integers(0, 256))
        cases.append(bytes(b[:int(rng.integers(0, len(b) + 1))]) if rng.random() < 0.3 else bytes(b))
    for data in cases:
        cur.write_bytes(data)
        try:
            jne.read_dat(cur)
        except jne.JneError:
            pass
        try:
            w = jne.DatWriter(cur, 4, 3, 77)
        except jne.JneError:
            continue  # other parameters in a damaged header: refused, as the reference does
        w.append([5], np.ones((1, 3)))
        w.finish()
        got_seeds, got = jne.read_dat(cur)[:2]
        assert got_seeds[-1] == 5 and np.array_equal(got[-1], np.ones(3))
