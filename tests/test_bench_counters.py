"""bench.py's counter summary picks the emit kernel instantiation that the
counter run timed (by family and grid) and scales each counter by the
copy kernel of known size (synthetic rocprofv3 CSVs, no GPU)."""
import csv
import json

import bench

HDR = ["Dispatch_Id", "Grid_Size", "Kernel_Name", "Counter_Name", "Counter_Value"]


def _csv(path, counter, rows):
    path.mkdir(parents=True)
    with open(path / "run_counter_collection.csv", "w", newline="") as fh:
        w = csv.writer(fh)
        w.writerow(HDR)
        for i, (name, grid, val) in enumerate(rows):
            w.writerow([i, grid, name, counter, val])


def test_summary_selects_timed_instantiation(tmp_path):
    f32_16 = "msh::(anonymous namespace)::k_emit_mc5<float, float, float, 16, true, 0>(msh::VolArgs)"
    f32_8 = "msh::(anonymous namespace)::k_emit_mc5<float, float, float, 8, true, 0>(msh::VolArgs)"
    f64_8 = "msh::(anonymous namespace)::k_emit_mc5<double, double, double, 8, true, 0>(msh::VolArgs)"
    h5mt = "msh::(anonymous namespace)::k_emit_h5<1, float, float, double, double, 4, true, 0>(msh::VolArgs)"
    h5sn8 = "msh::(anonymous namespace)::k_emit_h5<2, float, float, float, float, 8, true, 0>(msh::VolArgs)"
    h5sn4 = "msh::(anonymous namespace)::k_emit_h5<2, float, float, float, float, 4, true, 0>(msh::VolArgs)"
    copy = "__amd_rocclr_copyBuffer"
    fetch = [(f64_8, 1000, 100.0), (f32_16, 900, 40.0), (f32_16, 900, 42.0), (f32_8, 900, 99.0),
             (f32_16, 20, 1.0), (h5mt, 50, 7.0), (h5sn4, 50, 3.0), (h5sn8, 50, 9.0)] + \
        [(copy, 1, 524288.0)] * 4
    write = [(f64_8, 1000, 80.0), (f32_16, 900, 30.0), (f32_16, 900, 30.0), (f32_8, 900, 77.0),
             (f32_16, 20, 2.0), (h5mt, 50, 5.0), (h5sn4, 50, 1.0), (h5sn8, 50, 4.0)] + \
        [(copy, 1, 1048576.0)] * 4
    _csv(tmp_path / "f", "FETCH_SIZE", fetch)
    _csv(tmp_path / "w", "WRITE_SIZE", write)
    line = {"float64": {"family_id": 2, "algorithmic_bytes": 1000},
            "float32": {"family_id": 2, "algorithmic_bytes": 1000},
            "gyroid_256_f32": {"mt": {"family_id": 2, "algorithmic_bytes": 10},
                               "sn": {"family_id": 2, "algorithmic_bytes": 10},
                               "mc": {"family_id": 2, "algorithmic_bytes": 10}}}
    (tmp_path / "line.json").write_text("noise\n" + json.dumps(line) + "\n")
    out = bench.summarize_counters(tmp_path / "f", tmp_path / "w", tmp_path / "line.json",
                                   out_file=tmp_path / "t.json")
    assert out["scale"] == {"FETCH_SIZE": 2.0, "WRITE_SIZE": 1.0}
    w = out["workloads"]
    f32 = w["mc_value_noise_1024_float32"]
    assert f32["kernel"] == f32_16 and f32["grid"] == 900 and f32["dispatches"] == 2
    assert f32["fetch_bytes"] == int(41.0 * 1024 * 2) and f32["write_bytes"] == 30 * 1024
    assert w["mc_value_noise_1024_float64"]["kernel"] == f64_8
    assert w["mc_gyroid_256_float32"]["grid"] == 20  # the small launch of the same kernel
    assert w["mt_gyroid_256_float32"]["kernel"] == h5mt
    assert w["sn_gyroid_256_float32"]["kernel"] == h5sn8  # family 2: full-depth bricks
    # a summary from another family is not used for the timed kernel
    bench_t = json.loads((tmp_path / "t.json").read_text())
    assert bench_t["workloads"]["mc_value_noise_1024_float32"]["family_id"] == 2


def test_summary_splits_workloads_on_one_instantiation(tmp_path):
    # Float32 value noise and the Float32 sphere at 1024^3 run the same
    # instantiations at the same grid (each autotunes over the other's
    # family too): each workload takes the dispatches from its own field
    # generation to the next
    f32 = "msh::(anonymous namespace)::k_emit_mc5<float, float, float, 16, true, 0, 2>(msh::VolArgs)"
    f32h = "msh::(anonymous namespace)::k_emit_mc5<float, float, float, 8, true, 0, 1>(msh::VolArgs)"
    f64 = "msh::(anonymous namespace)::k_emit_mc5<double, double, double, 8, true, 0, 2>(msh::VolArgs)"
    gen = "msh::(anonymous namespace)::k_generate<float>(long)"
    copy = "__amd_rocclr_copyBuffer"

    def rows(vn, sph, kib_of_copy):
        r = [(gen, 7, 0.0), (f64, 1000, 1.0)]
        r += [(gen, 7, 0.0)] + [(f32h, 900, 99.0)] + [(f32, 900, vn)] * 3
        r += [(gen, 7, 0.0)] + [(f32, 900, 77.0)] + [(f32h, 900, sph)] * 2
        return r + [(gen, 7, 0.0)] + [(copy, 1, kib_of_copy)] * 4

    _csv(tmp_path / "f", "FETCH_SIZE", rows(40.0, 20.0, 524288.0))
    _csv(tmp_path / "w", "WRITE_SIZE", rows(30.0, 1.0, 1048576.0))
    line = {"float64": {"family_id": 2, "algorithmic_bytes": 1000},
            "float32": {"family_id": 2, "algorithmic_bytes": 1000},
            "sphere_1024_f32": {"family_id": 3, "algorithmic_bytes": 500}}
    (tmp_path / "line.json").write_text(json.dumps(line) + "\n")
    out = bench.summarize_counters(tmp_path / "f", tmp_path / "w", tmp_path / "line.json",
                                   out_file=tmp_path / "t.json")
    w = out["workloads"]
    vn, sp = w["mc_value_noise_1024_float32"], w["mc_sphere_1024_float32"]
    assert vn["dispatches"] == 3 and sp["dispatches"] == 2
    assert vn["fetch_bytes"] == 40 * 1024 * 2 and sp["fetch_bytes"] == 20 * 1024 * 2
    assert vn["write_bytes"] == 30 * 1024 and sp["write_bytes"] == 1 * 1024
