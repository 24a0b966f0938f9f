"""Extract the known-answer vectors of the reference's hot-path unit tests into JSON.

Run in a container that holds the reference sources:
    python tests/golden/extract_reference_kat.py /root/reference tests/golden/reference_kat.json

It reads the Catch2 test sources as text and keeps only data: for every TEST_CASE, the
parameters it sets, the chromosome, the LEF/barrier/rank/move inputs, the test hooks it calls
(in order) and the expected collisions, moves and ranks. The JSON it writes is the fixture the
tests replay against the oracle and the device step kernels; nothing here is executed from the
reference. Sources:
    test/units/simulation_cpu/simulation_complex_unit_test.cpp  (Simulation 001-012)
    test/units/simulation_cpu/simulation_simple_unit_test.cpp   (adjust moves, LEF-LEF, LEF-BAR,
                                                                 ranking)
    test/units/simulation_cpu/common.hpp:21 (DEFAULT_PRNG seed), :135-149 (init_config)
"""
import json
import re
import sys

DEFAULT_PRNG_SEED = 10556020843759504871
EVENTS = {"CHROM_BOUNDARY": 0x18, "LEF_BAR": 0x14, "LEF_LEF_PRIMARY": 0x12,
          "LEF_LEF_SECONDARY": 0x11}
HOOKS = ["test_adjust_and_clamp_moves", "test_process_collisions",
         "test_fix_secondary_lef_lef_collisions", "test_detect_units_at_chrom_boundaries",
         "test_detect_primary_lef_lef_collisions", "test_process_lef_lef_collisions",
         "test_detect_lef_bar_collisions", "test_correct_moves_for_lef_bar_collisions",
         "test_rank_lefs", "test_adjust_moves"]


def split_cases(text):
    out = []
    for m in re.finditer(r'TEST_CASE\("([^"]+)"', text):
        start = m.start()
        depth, i, began = 0, text.index("{", m.end()), False
        j = i
        while True:
            if text[j] == "{":
                depth += 1
            elif text[j] == "}":
                depth -= 1
                if depth == 0:
                    break
            j += 1
        out.append((m.group(1), text[i + 1:j]))
    return out


def strip_comments(s):
    s = re.sub(r"/\*.*?\*/", "", s, flags=re.S)
    return re.sub(r"//[^\n]*", "", s)


def parse_collision(tok):
    tok = tok.strip()
    m = re.fullmatch(r"CollisionT\{\s*\}", tok)
    if m:
        return [0, 0]
    m = re.fullmatch(r"CollisionT\{\s*(\d+)\s*,\s*(\w+)\s*\}", tok)
    return [int(m.group(1)), EVENTS[m.group(2)]]


def split_top(s):
    parts, depth, cur = [], 0, ""
    for ch in s:
        if ch in "({":
            depth += 1
        elif ch in ")}":
            depth -= 1
        if ch == "," and depth == 0:
            parts.append(cur)
            cur = ""
        else:
            cur += ch
    if cur.strip():
        parts.append(cur)
    return [p.strip() for p in parts if p.strip()]


def parse_case(name, body):
    body = strip_comments(body)
    case = {"name": name, "config": {}, "hooks": [], "arrays": {}, "rank_checks": []}
    m = re.search(r"init_config\(([^)]*)\)", body)
    if m:
        args = [float(a) for a in m.group(1).split(",")]
        case["config"].update(rev_speed=int(args[0]), fwd_speed=int(args[1]),
                              rev_speed_std=args[2] if len(args) > 2 else 0.0,
                              fwd_speed_std=args[3] if len(args) > 3 else 0.0,
                              p_bypass=0.0, pblock_major=1.0, pblock_minor=0.0)
    renames = {"lef_bar_major_collision_pblock": "pblock_major",
               "lef_bar_minor_collision_pblock": "pblock_minor",
               "probability_of_extrusion_unit_bypass": "p_bypass",
               "rev_extrusion_speed_std": "rev_speed_std",
               "fwd_extrusion_speed_std": "fwd_speed_std"}
    for m in re.finditer(r"\bc\.(\w+)\s*=\s*([^;]+);", body):
        key, val = m.group(1), m.group(2).strip()
        key = renames.get(key, key)
        try:
            case["config"][key] = float(val)
        except ValueError:
            case["config"][key] = val
    m = re.search(r'Chromosome chrom\{\s*0\s*,\s*"(\w+)"\s*,\s*(\d+)\s*,\s*(\d+)\s*,\s*(\d+)\s*\}', body)
    if m:
        case["chrom"] = dict(name=m.group(1), start=int(m.group(2)), end=int(m.group(3)),
                             size=int(m.group(4)))
    m = re.search(r'init_chromosome\(\s*"(\w+)"\s*,\s*(\d+)\s*(?:,\s*(\d+))?\s*(?:,\s*(\d+))?\s*\)', body)
    if m:
        size = int(m.group(2))
        start = int(m.group(3)) if m.group(3) else 0
        end = min(int(m.group(4)) if m.group(4) else size, size)
        case["chrom"] = dict(name=m.group(1), start=start, end=end, size=size)
    m = re.search(r"random::PRNG\((\d+)ULL\)", body)
    case["seed"] = int(m.group(1)) if m else DEFAULT_PRNG_SEED
    decls = re.finditer(r"std::(?:array<(\w+),\s*\w+>|vector<(\w+)>)\s+(\w+)\s*\{([^;]*)\};", body,
                        flags=re.S)
    for m in decls:
        typ, var, init = m.group(1) or m.group(2), m.group(3), m.group(4)
        items = split_top(init)
        if typ == "Lef":
            vals = []
            for it in items:
                mm = re.fullmatch(r"construct_lef\(\s*(\d+)\s*,\s*(\d+)\s*(?:,\s*(\d+))?\s*\)", it)
                vals.append([int(mm.group(1)), int(mm.group(2)), int(mm.group(3) or 0)])
        elif typ == "CollisionT":
            vals = [parse_collision(it) for it in items] if items else None
        else:
            vals = [int(it) for it in items] if items else None
        if vals is not None:
            case["arrays"][var] = {"type": typ, "values": vals}
    for m in re.finditer(r"(\w+)\[(\d+)\]\.release\(\)", body):
        arr = case["arrays"][m.group(1)]["values"]
        arr[int(m.group(2))] = None
    bm = re.search(r"construct_barriers(<\s*s\s*>)?\((.*?)\);", body, flags=re.S)
    if bm:
        inactive = bool(bm.group(1)) and "State::INACTIVE" in body
        bars = []
        for it in re.finditer(r"ExtrusionBarrier\{\s*(\d+)\s*,\s*([\d.]+)\s*,\s*([\d.]+)\s*,\s*'([+-])'\s*\}",
                              bm.group(2)):
            bars.append(dict(pos=int(it.group(1)), stp_active=float(it.group(2)),
                             stp_inactive=float(it.group(3)),
                             dir=1 if it.group(4) == "+" else 2,
                             state=0 if inactive else 1))
        case["barriers"] = bars
    elif "ExtrusionBarriers barriers{}" in body:
        case["barriers"] = []
    for m in re.finditer(r"(" + "|".join(HOOKS) + r")\(([^;]*)\);", body, flags=re.S):
        args = split_top(m.group(2))
        case["hooks"].append({"hook": m.group(1), "args": args})
    for m in re.finditer(r"CHECK\((\w+_ranks)\[(\d+)\]\s*==\s*(\d+)\)", body):
        case["rank_checks"].append([m.group(1), int(m.group(2)), int(m.group(3))])
    return case


def main(ref_root, out_path):
    cases = []
    for rel in ["test/units/simulation_cpu/simulation_complex_unit_test.cpp",
                "test/units/simulation_cpu/simulation_simple_unit_test.cpp"]:
        with open(f"{ref_root}/{rel}") as fh:
            text = fh.read()
        for name, body in split_cases(text):
            case = parse_case(name, body)
            case["source"] = rel
            cases.append(case)
    with open(out_path, "w") as fh:
        json.dump(cases, fh, indent=1)
    print(f"wrote {len(cases)} cases to {out_path}")


if __name__ == "__main__":
    main(sys.argv[1], sys.argv[2])
