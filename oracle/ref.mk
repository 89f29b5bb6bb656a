# CHECKER ONLY: compiles the reference's maths sources where they lie under
# $(REF) (no reference build system, no stand-in headers) plus ref_shim.cpp
# into _ref/libref_math.so, with the strict IEEE flags of the reference's
# unit-test build.  Outputs go to _ref/ only.
REF ?= /root/reference
CXX ?= g++
CXXFLAGS = -std=c++17 -O2 -fPIC -ffp-contract=off -fno-fast-math -w -I$(REF)/include
SRCS = $(REF)/src/math/bezier_utils.cpp $(REF)/src/math/mat2d_find_max_scale.cpp \
       $(REF)/src/math/mat2d.cpp ref_shim.cpp

_ref/libref_math.so: $(SRCS)
	mkdir -p _ref
	$(CXX) $(CXXFLAGS) -shared -o $@ $(SRCS)
