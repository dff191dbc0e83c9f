// placeholder, filled with the guiding-MLP oracle
