"""Every global name that a function of bench.py or __graft_entry__.py reads is defined: these
scripts run their main paths only on a GPU, so a misspelt or removed helper would otherwise first
show as a NameError there."""
import builtins
import importlib.util
import os
import symtable

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _undefined_globals(path):
    src = open(path).read()
    spec = importlib.util.spec_from_file_location(os.path.basename(path)[:-3] + "_names", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)  # module level only: main() runs under __name__ == "__main__"
    missing = []

    def walk(t):
        for ch in t.get_children():
            walk(ch)
        if t.get_type() == "module":
            return
        for s in t.get_symbols():
            if s.is_referenced() and s.is_global() and not s.is_assigned():
                n = s.get_name()
                if not hasattr(mod, n) and not hasattr(builtins, n):
                    missing.append(f"{t.get_name()}:{n}")

    walk(symtable.symtable(src, path, "exec"))
    return missing


@pytest.mark.parametrize("name", ["bench.py", "__graft_entry__.py"])
def test_no_undefined_globals(name):
    assert _undefined_globals(os.path.join(ROOT, name)) == []


def test_the_check_finds_a_missing_helper(tmp_path):
    p = tmp_path / "probe.py"
    p.write_text("def f():\n    return helper_that_is_not_there(1)\n\ndef g():\n    return len([])\n")
    assert _undefined_globals(str(p)) == ["f:helper_that_is_not_there"]


def _bench():
    spec = importlib.util.spec_from_file_location("bench_consts", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_measured_counters_match_the_sources_and_their_profiles():
    """Each counter reading bench.py reports was taken on the kernel sources as committed (their hash
    is the one recorded beside the reading), and each rocprof summary it cites exists and, where it
    quotes a source hash, quotes that one."""
    import re

    b = _bench()
    for (kernel, workload), m in b.MEASURED.items():
        assert b.src_sha(kernel) == m["src_sha"], (kernel, workload, "sources changed since the counters were read")
        for prof in m["profiles"]:
            path = os.path.join(ROOT, prof)
            assert os.path.exists(path), prof
            quoted = re.findall(r"source hash\s+`([0-9a-f]{16})`", open(path).read())
            assert all(q == m["src_sha"] for q in quoted), (prof, quoted, m["src_sha"])
