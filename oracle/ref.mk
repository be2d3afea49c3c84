# Builds oracle/_ref/libnepref.so from the reference's own standalone NEP CPU
# sources, where they lie in the reference tree (used only as a cross-check of
# the oracle's NEP restatement; never shipped or called by the product):
#   src/op/src/nep_cpu.cpp                  (fc, Chebyshev fn, s_lm, q, C3B/C4B/C5B)
#   src/op/src/cpu_calculate_nepneighbor.cpp (replica+MIC neighbor search, descriptor)
# cpu_calculate_nepneighbor.cpp includes <torch/extension.h>; the installed
# torch headers are used as they are (its functions take raw pointers only).
REF     ?= /root/reference
HERE    := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
OUT     := $(HERE)_ref/libnepref.so
PYTHON  ?= python3
TORCH_INC := $(shell $(PYTHON) -c "import torch.utils.cpp_extension as c; print(' '.join('-I'+p for p in c.include_paths()))")
PY_INC  := $(shell $(PYTHON) -c "import sysconfig; print(sysconfig.get_paths()['include'])")
SRCS    := $(REF)/src/op/src/cpu_calculate_nepneighbor.cpp $(REF)/src/op/src/nep_cpu.cpp

all: $(OUT)

$(OUT): $(SRCS)
	mkdir -p $(HERE)_ref
	g++ -O2 -fPIC -shared -std=c++17 $(TORCH_INC) -I$(PY_INC) $(SRCS) -o $@

clean:
	rm -f $(OUT)

.PHONY: all clean
