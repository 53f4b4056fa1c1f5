"""Counter-based traffic accounting (enterprise_amd.traffic): kernel-name
parsing, per-search sums and the copy-kernel calibration, on synthetic
rocprofv3 counter rows."""
import csv

import pytest

from enterprise_amd import traffic as T


def _write(path, rows, counter):
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
        for i, (name, val) in enumerate(rows):
            w.writerow([i, name, counter, val])


def test_kernel_base():
    assert T.kernel_base("void ent::k_level<256, false, false>(ent::Args, long long)") == "ent::k_level"
    assert T.kernel_base("ent::k_calib_copy(HIP_vector_type<int, 4u> const*, int*, long long)") == "ent::k_calib_copy"
    assert T.kernel_base("void ent::k_blevel_bu<256, 8>(ent::BArgs, long long)") == "ent::k_blevel_bu"


def test_per_search_plain_and_batch(tmp_path):
    rows = [("__amd_rocclr_fillBufferAligned", 5.0),
            ("void ent::k_init<256>(ent::Args, int, int)", 10.0),
            ("void ent::k_level<256, false, false>(ent::Args, long long)", 100.0),
            ("void ent::k_level<256, false, false>(ent::Args, long long)", 50.0),
            ("void ent::k_init<256>(ent::Args, int, int)", 10.0),
            ("void ent::k_level<256, false, false>(ent::Args, long long)", 30.0)]
    p = tmp_path / "c.csv"
    _write(p, rows, "FETCH_SIZE")
    got = T.read_counters(str(p), "FETCH_SIZE")
    assert len(got) == 6
    tot, n, by = T.per_search(got, "plain")
    assert (tot, n) == (200.0, 2)
    assert by == {"ent::k_init": 20.0, "ent::k_level": 180.0}
    brows = [("ent::k_binit_zero(ent::BArgs)", 1.0), ("ent::k_binit_src(ent::BArgs, int const*, ent::BInit)", 2.0),
             ("void ent::k_blevel<256, 8>(ent::BArgs, long long)", 3.0),
             ("void ent::k_blevel_bu<256, 8>(ent::BArgs, long long)", 4.0),
             ("void ent::k_level<256, false, false>(ent::Args, long long)", 99.0)]
    tot, n, _ = T.per_search(brows, "batch")
    assert (tot, n) == (10.0, 1)


def test_calib_factor():
    rows = [("ent::k_calib_copy(int4 const*, int4*, long long)", v) for v in (524300.0, 524298.0, 524298.0)]
    f, med = T.calib_factor(rows, 1 << 30)
    assert med == 524298.0 and f == pytest.approx(2.0, rel=1e-4)
    with pytest.raises(RuntimeError):
        T.calib_factor([("other", 1.0)], 1 << 30)


def test_committed_profiles_are_consistent():
    for name in ("plain_rmat22", "batch_rmat23"):
        t = T.load(T.os.path.join(T.ROOT, "profiles", f"traffic_{name}.json"))
        assert t is not None, name
        c = t["copy_kernel"]
        fetch = t["fetch_bytes_per_search"]
        assert t["traffic_bytes_per_search"] == pytest.approx(fetch + t["write_bytes_per_search"], abs=2)
        kib = sum(t["fetch_KiB_by_kernel_per_search"].values())
        assert fetch == pytest.approx(kib * 1024 * c["fetch_scale"], rel=1e-3)
        assert T.figure(name) == t["traffic_bytes_per_search"]
