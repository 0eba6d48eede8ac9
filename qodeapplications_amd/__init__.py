"""MI355X-native XR density build and Hamiltonian contraction."""
