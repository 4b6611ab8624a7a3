# Builds the reference's own Tecplot writer (ordered_tec.cpp) from the reference
# checkout, in place, with the driver plt_ref_driver.cpp:
#   make -C oracle -f ref.mk          -> oracle/_ref/plt_ref
# Only ordered_tec is buildable here: FluidCuda.cu is CUDA (no nvcc in this
# image), so the step itself is checked against the restatement in
# cpmp_oracle.cpp.
REF      ?= /root/reference
CXX      ?= g++
OUT      := _ref

all: $(OUT)/plt_ref

$(OUT)/plt_ref: plt_ref_driver.cpp $(REF)/ordered_tec.cpp $(REF)/ordered_tec.h
	@mkdir -p $(OUT)
	$(CXX) -O2 -std=c++11 -w -I$(REF) -o $@ plt_ref_driver.cpp $(REF)/ordered_tec.cpp

.PHONY: all
