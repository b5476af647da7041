"""Source rules of the HIP library, checked on the text of csrc/.

HIP refuses device (de)allocation while any stream of the process captures
a graph, and a copy on the legacy stream waits on (and fails against) a
capturing blocking stream.  The library therefore allocates, frees and
copies only through the wrappers in dim3_internal.h (dev_malloc, dev_free,
dev_malloc_async, dev_free_async, host_malloc, host_free, copy_wait), which
take the process-wide capture lock and keep copies on the caller's stream.
A raw call anywhere else in csrc/*.hip brings back the two failures seen
when a batch capture overlapped a server worker."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "joinproject-dim3_amd", "csrc")
RAW = re.compile(r"\b(hipMalloc|hipFree|hipMallocAsync|hipFreeAsync|hipHostMalloc|hipHostFree|hipMemcpy|hipMemcpy2D|"
                 r"hipMallocManaged|hipExtMallocWithFlags|hipHostRegister)\s*\(")


def code_of(path):
    src = open(path).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    src = re.sub(r'"(\\.|[^"\\])*"', '""', src)  # string literals (error messages)
    return re.sub(r"//[^\n]*", "", src)


def test_no_raw_allocation_or_legacy_copy_outside_the_wrappers():
    bad = []
    for name in sorted(os.listdir(CSRC)):
        if not name.endswith(".hip"):
            continue
        for i, line in enumerate(code_of(os.path.join(CSRC, name)).splitlines(), 1):
            if RAW.search(line):
                bad.append("%s:%d: %s" % (name, i, line.strip()))
    assert not bad, "raw HIP allocation / legacy-stream copy outside dim3_internal.h:\n" + "\n".join(bad)


def test_wrappers_take_the_capture_lock():
    src = code_of(os.path.join(CSRC, "dim3_internal.h"))
    for w in ("dev_malloc", "dev_free", "dev_malloc_async", "dev_free_async", "host_malloc", "host_free"):
        m = re.search(r"inline hipError_t %s\([^)]*\)\s*\{(.*?)\n\}" % w, src, flags=re.S)
        assert m and "capture_mutex()" in m.group(1), w
    m = re.search(r"inline hipError_t copy_wait\([^)]*\)\s*\{(.*?)\n\}", src, flags=re.S)
    assert m and "hipMemcpyAsync" in m.group(1) and "hipMemcpy(" not in m.group(1)
