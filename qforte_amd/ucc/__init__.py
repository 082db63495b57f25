"""UCC-VQE, ADAPT-VQE and UCC-PQE drivers."""
