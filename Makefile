# In-tree build with hipcc (gfx950) and g++. The targets mirror the reference's Makefile:
#   make            libddx.so with its signature-entry and shared-TT builds, the command lines
#                   (dds_gpu_amd/bin/, bin/sig/, bin/shared/) and the oracle
#   make gpu        the same: one library holds the CPU and the GPU solvers
#   make check      the CPU tests (pytest -m "not gpu"); make check-gpu runs the GPU tests
#   make clean
PYTHON ?= python3

all gpu:
	$(PYTHON) -c "import __graft_entry__ as g; g.build()"

check:
	$(PYTHON) -m pytest tests -q -m "not gpu"

check-gpu:
	$(PYTHON) -m pytest tests -q -m gpu

clean:
	rm -rf dds_gpu_amd/_build dds_gpu_amd/lib dds_gpu_amd/bin oracle/bin oracle/_ref

.PHONY: all gpu check check-gpu clean
