#! /bin/bash
# CK compile script: builds the engine library and the ck-bc-hip program with
# the package Makefile, then copies the program into CK's tmp directory.
set -e
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
make -s -C "$HERE/../.." libbcmi.so ck-bc-hip
cp -f "$HERE/../../libbcmi.so" .
cp -f "$HERE/../../ck-bc-hip" "./${CK_PROG_TARGET_EXE:-ck-bc-hip}"
