"""The rocprofv3 summariser that writes profiles/: per-dispatch means, copy calibration, name cleanup."""
import csv
import json
import os

from kaspan_amd import rocprof_summary as RS

FIELDS = ["Correlation_Id", "Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]


def write_pmc(d, rows):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "run_counter_collection.csv"), "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=FIELDS)
        w.writeheader()
        for r in rows:
            w.writerow(dict(zip(FIELDS, r)))


def test_pmc_means_and_calibration(tmp_path, capsys):
    k = "void kspn::k_expand<0, 8>(unsigned int const*, OpArgs)"
    # two dispatches of the kernel, the second with its counter split over two instances
    write_pmc(tmp_path / "f", [(1, 1, k, "FETCH_SIZE", 100.0), (2, 2, k, "FETCH_SIZE", 150.0),
                               (2, 2, k, "FETCH_SIZE", 50.0)])
    write_pmc(tmp_path / "w", [(1, 1, k, "WRITE_SIZE", 10.0), (2, 2, k, "WRITE_SIZE", 30.0)])
    copy = "void at::native::copy_kernel(int)"
    write_pmc(tmp_path / "cf", [(1, 1, copy, "FETCH_SIZE", 512.0)])  # 1 MiB copy read as 512 KiB: scale 2
    write_pmc(tmp_path / "cw", [(1, 1, copy, "WRITE_SIZE", 1024.0)])
    out = tmp_path / "t.json"
    RS.main(["pmc", "--fetch", str(tmp_path / "f"), "--write", str(tmp_path / "w"), "--copy-fetch",
             str(tmp_path / "cf"), "--copy-write", str(tmp_path / "cw"), "--copy-bytes", str(1 << 20),
             "--workload", "x", "--title", "t", "--command", "c", "--json", str(out)])
    data = json.load(open(out))
    assert abs(data["fetch_scale"] - 2.0) < 1e-12 and abs(data["write_scale"] - 1.0) < 1e-12
    e = data["kernels"]["kspn::k_expand<0, 8>"]
    assert e["launches"] == 2
    assert abs(e["fetch_kib"] - 150.0) < 1e-9 and abs(e["write_kib"] - 20.0) < 1e-9
    assert abs(e["bytes_per_launch"] - (150.0 * 1024 * 2 + 20.0 * 1024)) < 1e-6
    assert "k_expand<0, 8>" in capsys.readouterr().out


def test_kernel_names():
    assert RS.kernel_name("void kspn::k_pull<1, 2>(unsigned int, unsigned int const*)") == "kspn::k_pull<1, 2>"
    assert RS.kernel_name("kspn::k_part_split(unsigned int)") == "kspn::k_part_split"
