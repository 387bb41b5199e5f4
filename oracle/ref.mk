# Builds oracle/_ref/libcmref.so from the reference engine's own CPU sources,
# compiled where they lie (no copies), plus ref_shim.cpp.  Only used when the
# reference tree is present; TEST INFRASTRUCTURE ONLY.
REF ?= /root/reference/engine
SRC = $(REF)/src/physics/barnes_hut_solver.cpp $(REF)/src/physics/octree_builder.cpp \
      $(REF)/src/physics/octree_pool.cpp $(REF)/src/physics/particle_system.cpp \
      $(REF)/src/physics/post_newtonian.cpp \
      $(REF)/src/job/job_system.cpp $(REF)/src/job/job_queue.cpp $(REF)/src/core/error.cpp
CXXFLAGS = -O2 -std=c++20 -fPIC -ffp-contract=off -I$(REF)/include -w

_ref/libcmref.so: $(SRC) ref_shim.cpp
	@mkdir -p _ref
	g++ $(CXXFLAGS) -shared -o $@ ref_shim.cpp $(SRC) -lpthread

clean:
	rm -rf _ref

.PHONY: clean
