"""Drop-in solver classes of the dense smoke/fire step (the reference's
``backend`` package), running on libfluxhip."""
