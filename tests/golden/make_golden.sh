#!/bin/sh
# Regenerates the golden vectors: scenario inputs (Python), then the reference
# solver src/solver.js under node writes the expected outputs.  Development
# container only; the reference is read from /root/reference.
set -e
cd "$(dirname "$0")/../.."
python3 tests/golden/gen_inputs.py
node tests/golden/make_golden.mjs "$@"
