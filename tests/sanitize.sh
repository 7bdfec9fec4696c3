#!/bin/bash
# CPU sanitizer runs of the host code: librawspec's C host layer (API
# semantics, RAW header parser, filterbank writer/reader), the rawspec
# program and its threaded ingest (through the host-only harness).
#
#   bash tests/sanitize.sh            # both passes
#   bash tests/sanitize.sh address    # ASan + UBSan only
#   bash tests/sanitize.sh thread     # TSan only
#
# The instrumented library and program are built with gcc's sanitizers under
# rawspec_amd/_san/<kind>/ (the HIP device layer is compiled as usual; only
# gcc-compiled host C is instrumented).  Every finding aborts its process, so a
# finding fails the test that ran it.  No GPU is needed.
set -euo pipefail
cd "$(dirname "$0")/.."
passes=${1:-"address thread"}
for kind in $passes; do
  case $kind in
    address) san=address,undefined ;;
    thread) san=thread ;;
    *) echo "unknown pass $kind" >&2; exit 2 ;;
  esac
  echo "== $san"
  python -m rawspec_amd.build --sanitize "$san" > /dev/null
  export RAWSPEC_SANITIZE=$san
  export RAWSPEC_AMD_CLI=$PWD/rawspec_amd/_san/$kind/bin/rawspec
  if [ "$kind" = address ]; then
    # the Python process loads the instrumented library, so the ASan runtime
    # is preloaded; leak checking is off (the interpreter keeps its objects)
    RAWSPEC_AMD_LIB=$PWD/rawspec_amd/_san/$kind/librawspec.so \
    LD_PRELOAD=$(gcc -print-file-name=libasan.so) \
    ASAN_OPTIONS=detect_leaks=0:alloc_dealloc_mismatch=0:detect_odr_violation=0 \
    UBSAN_OPTIONS=print_stacktrace=1 \
      python -m pytest -q -p no:cacheprovider -m "not gpu" \
        tests/test_ingest_host.py tests/test_cli_host.py tests/test_host_formats.py tests/test_abi.py
  else
    # TSan runs instrumented processes only: the ingest harness and the
    # program (the Python process keeps the normal library)
    TSAN_OPTIONS=halt_on_error=1:second_deadlock_stack=1 \
      python -m pytest -q -p no:cacheprovider -m "not gpu" tests/test_ingest_host.py tests/test_cli_host.py
  fi
done
