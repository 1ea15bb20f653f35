#!/bin/bash
# Regenerates philox_rocrand.json from rocRAND's host engine (host-only build).
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
TMP=$(mktemp -d)
hipcc --cuda-host-only -O1 -o "$TMP/kat" "$HERE/rocrand_philox_vectors.cpp"
"$TMP/kat" > "$HERE/philox_rocrand.json"
rm -rf "$TMP"
