"""The batched renderer's frame planner (pupiloptixlab_amd/csrc/pt_batch_plan.h),
compiled on the host with tests/plan_sim.cpp, which walks a render's launches
the way pt_batch_render does and applies given gate decisions.

Checked: every frame of every view is rendered exactly once and in order; a
launch stays within its path capacity and holds either training frames only
or none; training frames go round-robin over the views; the gate's
evaluation frames are [T/4, T/2) of a view training on T >= 16 frames,
alternately guided and plain, with one decision frame at T/2 - 1; a view the
gate switches off trains on no frame from T/2 on, a view that keeps the
guide trains to T and restarts its accumulation at its first frame after T;
segment bookkeeping (first entry, frame counts, pixel offsets, sample counts)
matches the entries."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TRAIN, PLAIN, EVAL, DECIDE = 0x80000000, 0x40000000, 0x20000000, 0x10000000


@pytest.fixture(scope="module")
def sim(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("plan") / "plan_sim")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-I", os.path.join(ROOT, "pupiloptixlab_amd", "csrc"),
                    os.path.join(ROOT, "tests", "plan_sim.cpp"), "-o", exe], check=True)
    return exe


def run(sim, capacity, train_capacity, gate, views):
    """views: (width, height, spp, sample_cnt, random_seed, T, off)."""
    text = f"{capacity} {train_capacity} {int(gate)} {len(views)}\n" + \
        "\n".join(" ".join(str(x) for x in v) for v in views) + "\n"
    out = subprocess.run([sim], input=text, capture_output=True, text=True, check=True).stdout
    launches = []
    for line in out.splitlines():
        f = line.split()
        vals = [int(x) for x in f[1:]]
        if f[0] == "L":
            launches.append({"total": vals[0], "train_frames": vals[1], "n_e": vals[2], "n_s": vals[3],
                             "e": [], "s": []})
        elif f[0] == "E":
            launches[-1]["e"].append(dict(zip(("view", "flags", "frame", "path_begin", "pixels"), vals)))
        else:
            launches[-1]["s"].append(dict(zip(("view", "first", "frames", "pix_begin", "cnt0", "restart",
                                               "restart_cnt0"), vals)))
    return launches


def check_common(launches, views, capacity, train_capacity):
    seen = [[] for _ in views]
    for l in launches:
        assert l["n_e"] == len(l["e"]) and l["n_s"] == len(l["s"])
        trains = [bool(e["flags"] & TRAIN) for e in l["e"]]
        assert all(trains) or not any(trains)  # training frames only, or none
        assert l["train_frames"] == sum(trains)
        cap = min(capacity, train_capacity) if any(trains) else capacity
        assert l["total"] <= cap or len(l["e"]) == 1
        path = 0
        for e in l["e"]:
            w, h = views[e["view"]][:2]
            assert e["path_begin"] == path and e["pixels"] == w * h
            path += w * h
            seen[e["view"]].append(e["frame"])
        assert path == l["total"]
        pix = 0
        for s in l["s"]:
            w, h, spp, cnt = views[s["view"]][:4]
            ents = l["e"][s["first"]:s["first"] + s["frames"]]
            assert ents and all(e["view"] == s["view"] for e in ents)
            assert [e["frame"] for e in ents] == list(range(ents[0]["frame"], ents[0]["frame"] + len(ents)))
            assert s["cnt0"] == cnt + ents[0]["frame"]
            assert s["pix_begin"] == pix
            pix += w * h
        assert sum(s["frames"] for s in l["s"]) == len(l["e"])
        assert len({s["view"] for s in l["s"]}) == len(l["s"])  # one segment per view and launch
    for v, frames in enumerate(seen):
        assert frames == list(range(views[v][2]))  # every frame once, in order


def test_plain_batch_packs_all_frames(sim):
    views = [(64, 64, 16, 0, 0, 0, 0), (32, 16, 4, 3, 100, 0, 0), (100, 50, 64, 0, 7, 0, 0)]
    launches = run(sim, 40000, 0, False, views)
    check_common(launches, views, 40000, 0)
    assert all(l["train_frames"] == 0 and not any(s["restart"] for s in l["s"]) for l in launches)
    assert all(e["flags"] == 0 for l in launches for e in l["e"])


def test_training_frames_round_robin(sim):
    views = [(64, 64, 40, 0, 0, 8, 0), (64, 64, 20, 0, 0, 8, 0), (32, 32, 12, 0, 0, 8, 0)]
    launches = run(sim, 1 << 20, 3 * 64 * 64, False, views)
    check_common(launches, views, 1 << 20, 3 * 64 * 64)
    order = [(e["view"], e["frame"]) for l in launches for e in l["e"] if e["flags"] & TRAIN]
    assert sorted(order) == sorted((v, f) for v in range(3) for f in range(8))
    # round r holds frame r of every view still training before frame r + 1 of any
    rounds = [f for _, f in order]
    assert rounds == sorted(rounds)
    # training launches come first, then the rest view by view
    first_plain = next(i for i, l in enumerate(launches) if l["train_frames"] == 0)
    assert all(l["train_frames"] == 0 for l in launches[first_plain:])
    rest = [e["view"] for l in launches[first_plain:] for e in l["e"]]
    assert rest == sorted(rest)


@pytest.mark.parametrize("off", [(0, 0, 0, 0), (1, 0, 1, 0), (1, 1, 1, 1), (0, 1, 0, 1)])
def test_gate_windows_decisions_and_restart(sim, off):
    # T = 32 (eval 8..15), T = 16 (eval 4..7, all frames training), T = 8
    # (not gated), T = 32 with spp 256
    views = [(48, 32, 64, 0, 0, 32, off[0]), (48, 32, 16, 0, 1000, 16, off[1]), (32, 32, 40, 5, 0, 8, off[2]),
             (64, 48, 256, 0, 77, 32, off[3])]
    cap, tcap = 64 * 48 * 6, 48 * 32 * 2
    launches = run(sim, cap, tcap, True, views)
    check_common(launches, views, cap, tcap)
    for v, (w, h, spp, cnt, seed, T, o) in enumerate(views):
        ents = [e for l in launches for e in l["e"] if e["view"] == v]
        eligible = T // 2 - T // 4 >= 4
        ev = [e["frame"] for e in ents if e["flags"] & EVAL]
        if eligible:
            assert ev == list(range(T // 4, T // 2))
            assert [bool(e["flags"] & PLAIN) for e in ents if e["flags"] & EVAL] == \
                [bool(i & 1) for i in range(len(ev))]
            assert [e["frame"] for e in ents if e["flags"] & DECIDE] == [T // 2 - 1]
            assert all(e["flags"] & TRAIN for e in ents if e["flags"] & EVAL)
        else:
            assert not ev and not any(e["flags"] & (DECIDE | PLAIN) for e in ents)
        gated = eligible and o
        train_end = T // 2 if gated else T
        assert [e["frame"] for e in ents if e["flags"] & TRAIN] == list(range(train_end))
        # restart: segments whose first frame is at or after T (the device
        # ignores it for a view whose guide is off)
        for l in launches:
            for s in l["s"]:
                if s["view"] != v:
                    continue
                f0 = l["e"][s["first"]]["frame"]
                assert s["restart"] == int(f0 >= T and T >= 4)
                if s["restart"]:
                    assert s["restart_cnt0"] == f0 - T


def test_gate_off_view_leaves_training_early(sim):
    # two views at T = 32; view 1 is switched off after frame 15: its frames
    # 16.. are plain and come after all training launches
    views = [(64, 64, 64, 0, 0, 32, 0), (64, 64, 64, 0, 0, 32, 1)]
    launches = run(sim, 64 * 64 * 8, 64 * 64 * 2, True, views)
    check_common(launches, views, 64 * 64 * 8, 64 * 64 * 2)
    train = [(e["view"], e["frame"]) for l in launches for e in l["e"] if e["flags"] & TRAIN]
    assert [f for v, f in train if v == 1] == list(range(16))
    assert [f for v, f in train if v == 0] == list(range(32))
    last_train = max(i for i, l in enumerate(launches) if l["train_frames"])
    assert all(l["train_frames"] == 0 for l in launches[last_train + 1:])
