#!/bin/bash
# Build tests/cpp/sanitized_host with AddressSanitizer + UndefinedBehaviorSanitizer: host code
# only (the oracle in C, the basis enumerator and the step search in C++, the TSV reader of
# include/dynamics.h), compiled by gcc/g++ with no GPU code and no OpenMP.
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
ROOT="$(cd "$HERE/../.." && pwd)"
OUT="${1:-$HERE/sanitized_host}"
TMP="$(mktemp -d)"
trap 'rm -rf "$TMP"' EXIT
SAN="-fsanitize=address,undefined -fno-sanitize-recover=all -fno-omit-frame-pointer -g -O1"
gcc $SAN -Wno-unknown-pragmas -c "$ROOT/oracle/le_oracle.c" -o "$TMP/le_oracle.o"
g++ $SAN -std=c++17 -c "$ROOT/laughlinedge_amd/csrc/le_basis.cpp" -o "$TMP/le_basis.o"
g++ $SAN -std=c++17 -c "$ROOT/laughlinedge_amd/csrc/le_tune.cpp" -o "$TMP/le_tune.o"
g++ $SAN -std=c++17 -I "$ROOT/include" -c "$HERE/sanitized_host.cpp" -o "$TMP/main.o"
g++ $SAN -o "$OUT" "$TMP/main.o" "$TMP/le_oracle.o" "$TMP/le_basis.o" "$TMP/le_tune.o" -lm
