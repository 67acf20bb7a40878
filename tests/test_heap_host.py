"""Host-only checks of the symmetric offset allocator (osc_heap.c), built
with gcc from the library's own source: no device needed.

The reference's dlmalloc returns NULL for a request larger than the heap
(src/api/shmalloc.c:47-61); the allocator must do the same for sizes near
SIZE_MAX instead of wrapping its rounding or its fit test."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "osss-ucx_amd", "csrc", "host")


def test_heap_rejects_wrapping_sizes(tmp_path):
    exe = str(tmp_path / "heap_alloc")
    r = subprocess.run(
        ["gcc", "-std=gnu11", "-O1", "-Wall", "-Wextra", "-Werror",
         "-Wno-unused-parameter", "-fsanitize=undefined",
         "-fno-sanitize-recover=all", "-I", HOST,
         os.path.join(ROOT, "tests", "c", "heap_alloc.c"),
         os.path.join(HOST, "osc_heap.c"), "-o", exe],
        capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stderr
