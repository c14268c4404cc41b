#!/bin/sh
# Regenerates tests/golden/philox_kat.json from PyTorch's Philox source header.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
TORCH_INC=$(python3 -c "import os, torch; print(os.path.join(os.path.dirname(torch.__file__), 'include'))")
g++ -std=c++17 -O1 -I"$TORCH_INC" "$HERE/gen_philox_kat.cpp" -o /tmp/gen_philox_kat
/tmp/gen_philox_kat > "$HERE/philox_kat.json"
