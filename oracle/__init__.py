"""CPU oracle for the grid tensor-network amplitude path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (``qflex_amd``) imports,
links or executes this package.  Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker
(or, for the baseline, as the CPU implementation being timed).

It is a clean-room numpy restatement of the reference qFlex CPU path:

* ``tensor_ops``   -- Tensor::reorder / multiply / bundle / project
                      (reference src/tensor.cpp)
* ``formats``      -- circuit / grid / ordering text formats
                      (reference src/circuit.cpp, src/grid.cpp, src/ordering.cpp)
* ``network``      -- gate tables, 3D network, 3D->2D flatten
                      (reference src/read_circuit.cpp)
* ``contraction``  -- ordering parser, ValidateOrdering, ContractionData,
                      get_output_states, EvaluateCircuit
                      (reference src/contraction_utils.cpp, src/evaluate_circuit.cpp)
* ``statevector``  -- an independent FP64 state-vector simulator built from
                      textbook gate matrices; it pins the tensor-network path.

Every tensor-network function runs in either complex128 (the "complex128
oracle") or complex64 (host reorders + numpy/OpenBLAS cgemm: the CPU
baseline, the same algorithm the reference runs).

Pinning (what checks each function; see tests/test_oracle_*.py):
  reorder      -- reference known answer [0,4,1,5,2,6,3,7]
                  (tests/src/tensor_test.cpp:234-266) and a brute-force
                  bit-permutation (tensor_test.cpp:44-74) on 10/12-index cases
  multiply     -- known answer {20,14,100,78} (tensor_test.cpp:496-530) and
                  a closed-form einsum on random complex tensors
  ContractGrid -- SimpleInitializeData -> [1, 0]
                  (tests/src/contraction_utils_test.cpp:244-301)
  flatten      -- CondenseToGrid index orders (tests/src/read_circuit_test.cpp:281-345)
  EvaluateCircuit -- SimpleCircuit known answer (tests/src/evaluate_circuit_test.cpp:185-207),
                  and the FP64 state vector on seeded random circuits (<= 20 qubits),
                  and sum |amp|^2 == 1 over all bitstrings on tiny grids
  get_output_states -- tests/src/evaluate_circuit_test.cpp:51-97
"""
