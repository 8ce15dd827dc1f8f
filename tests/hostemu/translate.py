"""Translates one of the library's .hip sources into host C++ over shim.h (test tooling).

Every `kernel<<<grid, block, shmem, stream>>>(args);` becomes a launch_cpu call that runs the
kernel body for each thread of the grid; kernels that use cross-lane operations are launched
as wave kernels (one fiber per GPU thread).
"""
import re
import sys

WAVE_KERNELS = {"count_centres", "write_centres", "dfs_replay", "cr_histogram", "scan_reduce<T>", "scan_down<T>",
                "prune_centres", "emit_centres", "cover_tiles", "count_equal_u8", "far_prefix_max",
                "relax_blocks", "find_start", "cavity_totals", "max_run", "partition_active", "next_min_key",
                "edt_pass_y", "edt_pass_z", "owned_counts",
                "ws_resolve", "ws_count_parts", "ws_scan_parts", "ws_append", "set_union_bitmap", "below_min", "length_histogram",
                "plateau_edges", "reach_sizes_lds", "list_counts_lds", "list_fill_lds"}


def split_args(s):
    s = s.replace("->", "\x01")
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "(<[":
            depth += 1
        if ch in ")>]":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur)
            cur = ""
        else:
            cur += ch
    out.append(cur)
    return [o.replace("\x01", "->") for o in out]


def launch(m):
    # The arguments are evaluated once, before the launch, as a real launch does.
    name, cfg, args = m.group(1), split_args(m.group(2)), m.group(3)
    wave = "true" if name in WAVE_KERNELS or name.split("<")[0] in WAVE_KERNELS else "false"
    shmem = cfg[2].strip() if len(cfg) > 2 else "0"
    return (f"{{ auto hostemu_args = std::make_tuple({args}); "
            f"launch_cpu({wave}, Dim3({cfg[0].strip()}), Dim3({cfg[1].strip()}), [&] {{ "
            f"std::apply([&](auto... a) {{ {name}(a...); }}, hostemu_args); }}, (size_t)({shmem})); }}")


def translate(text):
    text = text.replace('#include "pma_internal.h"', '#include "shim.h"')
    text = re.sub(r"#include <rocprim/.*?>", "", text)
    text = re.sub(r"extern __shared__ (\w+(?: \w+)*?) (\w+)\[\];", r"\1* \2 = (\1*)hostemu_dyn_shared();", text)
    return re.sub(r"(\w+(?:<[\w, ]*>)?)<<<(.*?)>>>\((.*?)\);", launch, text, flags=re.S)


if __name__ == "__main__":
    src, dst = sys.argv[1], sys.argv[2]
    with open(src) as f, open(dst, "w") as g:
        g.write(f"// translated from {src}\n" + translate(f.read()))
