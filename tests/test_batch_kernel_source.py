"""Source invariants of the batch engine's step kernels (csrc/ed_batch.hip), checked on the CPU.

The convergence check runs on a side stream beside the next chunk's step kernels and switches
copies off (their `active` flag, their group's running count `gact`) while those kernels run.
A workgroup whose waves decided separately whether to leave could lose a wave that another wave
still needs: in k_g_up, wave r writes row r's hop descriptors and their count to LDS and every
wave gathers with all of them, so a wave left behind would gather from stale descriptors.  Every
leave-or-stay decision on these flags in a kernel of more than one wave therefore goes through
group_running (thread 0 reads, the workgroup follows); a per-lane `active` read may only gate
the lane's own work.
"""
import os
import re

SRC = os.path.join(os.path.dirname(__file__), "..", "exact_diagonalization_for_fermionic_systems_with_gpu_amd", "csrc",
                   "ed_batch.hip")


def _kernels(text):
    """name -> (launch bound, body) of every __global__ function"""
    out = {}
    for m in re.finditer(r"__global__\s+(?:__launch_bounds__\((\d+)\)\s+)?void\s+(\w+)\s*\(", text):
        i = text.index("{", m.end())
        depth, j = 0, i
        while True:
            if text[j] == "{":
                depth += 1
            elif text[j] == "}":
                depth -= 1
                if depth == 0:
                    break
            j += 1
        out[m.group(2)] = (int(m.group(1)) if m.group(1) else None, text[i:j + 1])
    return out


def test_group_count_read_only_through_group_running():
    text = open(SRC).read()
    reads = [ln for ln in text.splitlines() if "gact[" in ln and "atomicSub" not in ln and "pgact" not in ln]
    assert len(reads) == 1 and "*s_flag = gact[grp]" in reads[0], reads


def test_step_kernels_leave_as_a_workgroup():
    text = open(SRC).read()
    ks = _kernels(text)
    for name in ("k_g_up", "k_g_dn", "k_g_update", "k_g_finalize"):
        bound, body = ks[name]
        assert bound is not None and bound > 64, name  # more than one wave per workgroup
        assert "group_running(" in body, name
        # no early return on a flag read by each thread itself
        assert not re.search(r"if\s*\([^)]*active\[[^)]*\)\s*return", body), name
        assert "gact[" not in body, name


def test_check_kernels_are_one_wave():
    ks = _kernels(open(SRC).read())
    for name in ("k_g_check", "k_g_check_k"):
        assert ks[name][0] == 64, name  # their own early return on active[c] is uniform
