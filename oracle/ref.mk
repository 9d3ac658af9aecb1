# Builds the reference solver itself from its sources where they lie
# (REF, read-only) into oracle/_ref/tensum.  Test infrastructure only: the
# binary writes the fixtures under tests/golden/ that pin the oracle; nothing
# in the product loads it.
#
# The reference needs MPI (every translation unit includes <mpi.h> through
# src/mpi_utils.h).  The image's MPICH 3.3.2 under /opt/conda is linked as it
# is: its shared libraries are reached through symlinks in _ref/lib so that
# the binary's run path does not also pull in conda's older libstdc++.  The
# conda compiler wrapper (mpicxx) is not used: it names a compiler the image
# lacks.  Usage: make -f ref.mk [-j8]
REF ?= /root/reference
CONDA ?= /opt/conda
OUT := _ref
LIB := $(OUT)/lib
CXX := g++
CXXFLAGS := -std=c++11 -O3 -w -DMPICH_SKIP_MPICXX -I$(CONDA)/include -I$(REF)/extern/fparser -I$(REF)/src

SRCS := $(wildcard $(REF)/src/*.cc)
OBJS := $(patsubst $(REF)/src/%.cc,$(OUT)/obj/%.o,$(SRCS)) $(OUT)/obj/fparser.o

# The reference's partitioner (grid_gen/, no MPI): gmsh-2.2 mesh with
# partition and ghost tags -> PARTITION/part_file_<p>.dat for multi-rank runs
# of the reference (the CPU baseline on several cores).  Built without
# optimisation: grid_gen/gmsh_split.cc:84-92 reads each element's tags into
# tags[1..ntags] of an ntags-long stack array, one past its end (an
# AddressSanitizer build reports the overflow there); at -O2 that write
# corrupts the stack and the tool crashes, at -O0 it runs.
GP_SRCS := $(wildcard $(REF)/grid_gen/*.cc)
GP_OBJS := $(patsubst $(REF)/grid_gen/%.cc,$(OUT)/gp_obj/%.o,$(GP_SRCS))

all: $(OUT)/tensum $(OUT)/grid_part

$(LIB)/.links:
	mkdir -p $(LIB) $(OUT)/obj
	ln -sf $(CONDA)/lib/libmpi.so.12 $(LIB)/libmpi.so.12
	ln -sf $(CONDA)/lib/libmpi.so.12 $(LIB)/libmpi.so
	ln -sf $(CONDA)/lib/libgfortran.so.4 $(LIB)/libgfortran.so.4
	ln -sf $(CONDA)/lib/libquadmath.so.0 $(LIB)/libquadmath.so.0
	touch $@

$(OUT)/obj/%.o: $(REF)/src/%.cc | $(LIB)/.links
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/obj/fparser.o: $(REF)/extern/fparser/fparser.cc | $(LIB)/.links
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/gp_obj/%.o: $(REF)/grid_gen/%.cc
	mkdir -p $(OUT)/gp_obj
	$(CXX) -std=c++11 -O0 -w -I$(REF)/grid_gen -c $< -o $@

$(OUT)/grid_part: $(GP_OBJS)
	$(CXX) -o $@ $(GP_OBJS)

$(OUT)/tensum: $(OBJS)
	$(CXX) -o $@ $(OBJS) -L$(abspath $(LIB)) -lmpi -ldl -Wl,-rpath,$(abspath $(LIB))

clean:
	rm -rf $(OUT)
.PHONY: all clean
