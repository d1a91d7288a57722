# ref.mk — builds the reference dnaln from its own sources (read-only checkout) into oracle/_ref/.
#
#   make -f oracle/ref.mk REF=/root/reference
#
# Outputs (all under oracle/_ref/, which git ignores): the `dnaln` CLI and libdnaln_ref.so
# (the reference objects plus oracle/ref_shim.cpp). The reference's own Makefile and build tree
# are not used; its sources are compiled in place with g++ and linked against the system zlib.
# -march=x86-64-v3 selects the reference's AVX2 x16 kernel and runs on any AVX2 host.
REF      ?= /root/reference
HERE     := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
OUT      := $(HERE)_ref
CXX      ?= g++
CXXFLAGS := -std=c++17 -O3 -DNDEBUG -march=x86-64-v3 -fPIC -w
NAMES    := common seqio reference minimizer index chain align align_simd mapper simulate evaluate \
            groundtruth bench demo
OBJS     := $(NAMES:%=$(OUT)/obj/%.o)

all: $(OUT)/dnaln $(OUT)/libdnaln_ref.so

$(OUT)/obj/%.o: $(REF)/src/%.cpp | $(OUT)/obj
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/obj/shim.o: $(HERE)ref_shim.cpp | $(OUT)/obj
	$(CXX) $(CXXFLAGS) -I$(REF)/src -c $< -o $@

$(OUT)/dnaln: $(OBJS) $(OUT)/obj/main.o
	$(CXX) $(CXXFLAGS) $^ -o $@ -lz

$(OUT)/libdnaln_ref.so: $(OBJS) $(OUT)/obj/shim.o
	$(CXX) $(CXXFLAGS) -shared $^ -o $@ -lz

$(OUT)/obj:
	mkdir -p $@

.PHONY: all
