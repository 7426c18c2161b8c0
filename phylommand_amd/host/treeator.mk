# Builds phylommand's treeator from its own, unmodified src/treeator.cpp with
# the likelihood classes (simpleML, sub_model, rate_model) taken from
# phylommand_amd/host instead of src/simpleML.cpp, src/sub_model.cpp and
# src/rate_model.cpp, linked to libphylommand_host.so and libphylo_hip.so.
#
#   make -f phylommand_amd/host/treeator.mk REF=/path/to/phylommand/src OUT=build_dir
#
# How the swap works:
#  - treeator.cpp includes "simpleML.h", "sub_model.h" and "rate_model.h"
#    with quotes, which would find src/'s own headers first.  -I- turns off
#    that search of the including file's directory: quoted includes look in
#    phylommand_amd/host (before -I-) and then in src/ (after it), so those
#    three come from this project and every other header (tree.h,
#    character_vector.h, matrix_parser.h, ...) from phylommand.
#  - PHYLO_AMD_AS_REFERENCE makes the host headers declare the global names
#    treeator uses (simpleML, sub_model, rate_model).
#  - -std=c++14: libstdc++'s C++17 parallel-algorithm headers include their
#    own files with quotes, which -I- would hide.
#  - phylommand's tree.o still needs its own rate_model.o (tree skyline
#    scaling) and marth.o (tree simulation); neither is on the likelihood
#    path, which goes through phylommand_amd's classes.
REF ?= /root/reference/src
OUT ?= build_treeator
HOST := $(patsubst %/,%,$(dir $(abspath $(lastword $(MAKEFILE_LIST)))))
PKG := $(patsubst %/,%,$(dir $(HOST)))
ROOT := $(patsubst %/,%,$(dir $(PKG)))
CXX ?= g++
REF_OBJS := tree string_tree nj_tree matrix_parser file_parser argv_parser rate_model marth

.PHONY: all clean
all: $(OUT)/treeator

$(OUT)/treeator.o: $(REF)/treeator.cpp $(HOST)/simpleML.h $(HOST)/sub_model.h $(HOST)/rate_model.h $(HOST)/phylo_tree.h
	mkdir -p $(OUT)
	$(CXX) -std=c++14 -O2 -w -DPHYLO_AMD_AS_REFERENCE -I$(HOST) -I- -I$(REF) -I$(ROOT)/include -c -o $@ $<

$(OUT)/ref_%.o: $(REF)/%.cpp
	mkdir -p $(OUT)
	$(CXX) -std=c++11 -O2 -w -I$(REF) -c -o $@ $<

$(OUT)/ref_marth.o: $(REF)/marth/marth.cpp
	mkdir -p $(OUT)
	$(CXX) -std=c++11 -O2 -w -I$(REF) -c -o $@ $<

$(OUT)/treeator: $(OUT)/treeator.o $(addprefix $(OUT)/ref_,$(addsuffix .o,$(REF_OBJS)))
	$(CXX) -o $@ $^ -L$(PKG) -lphylommand_host -lphylo_hip -Wl,-rpath,$(PKG)

clean:
	rm -f $(OUT)/treeator $(OUT)/*.o
