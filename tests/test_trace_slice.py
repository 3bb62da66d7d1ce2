"""CPU test of the kernel-trace slicing used for the per-iteration solve profile."""
import csv

from terraneo_amd.kernel_timing import slice_trace


def _trace(path):
    rows = []
    t = 1000

    def add(name, dur, gap=10):
        nonlocal t
        t += gap
        rows.append({"Kernel_Name": name, "Start_Timestamp": t, "End_Timestamp": t + dur})
        t += dur

    for solve in range(2):
        add("radix_sort_onesweep", 50)
        add("radix_sort_upsweep", 50)
        for _ in range(2):                       # two iterations
            add("k_eps_march<double>", 4000 if solve == 0 else 2000)
            add("k_dot_owned", 100, gap=500)
        add("radix_sort_onesweep", 50, gap=10)
        for _ in range(4):                       # set-up work between the solves
            add("k_setup", 10)
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        w.writeheader()
        for r in rows[::-1]:                     # unsorted input
            w.writerow(r)


def test_slice_trace(tmp_path, capsys):
    p = tmp_path / "run_kernel_trace.csv"
    _trace(p)
    slice_trace(str(p), reps=2)
    out = capsys.readouterr().out
    blocks = out.split("== ")[1:]
    assert [b.split(":")[0] for b in blocks] == ["f64", "mixed"]
    # f64 segment: per iteration 4.0 us march + 0.1 us dot, idle 0.01 + 0.5 us
    assert "4 kernels" in blocks[0]
    assert "kernel-busy 0.00 ms" in blocks[0]
    line = [l for l in blocks[0].splitlines() if "k_eps_march" in l][0]
    assert line.split()[0] == "0.004" and line.split()[2] == "1.0"
    line = [l for l in blocks[1].splitlines() if "k_eps_march" in l][0]
    assert line.split()[0] == "0.002"
    gap = [l for l in blocks[0].splitlines() if l.strip().endswith("k_eps_march<double>")][-1]
    assert gap.split()[0] == "0.001"             # 2 x 500 ns after the march, over 2 iterations
