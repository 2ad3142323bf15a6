def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device")
