"""The phantom generator's statistical checks (tests/test_phantom_generate.py:
region weights and kinds, non-additive rejection and additive overlap, the
detection rule and noise widths, the images and tan-bins map, the expected
image of detected emissions) on events generated on the device, so that the
device stream is held to the reference's rules directly and not only to its
host twin."""
import pytest

import test_phantom_generate as tpg

pytestmark = pytest.mark.gpu

CHECKS = [tpg.test_region_weights_and_kinds,
          tpg.test_non_additive_rejection_and_additive_overlap,
          tpg.test_detection_rule_and_noise,
          tpg.test_images_and_tan_bins,
          tpg.test_expected_image_matches_generated_emissions]


@pytest.mark.parametrize("check", CHECKS, ids=[c.__name__[5:] for c in CHECKS])
def test_device_events_follow_the_rules(gpu_pkg, monkeypatch, check):
    print(f"case: {check.__name__} on device 0", flush=True)
    monkeypatch.setattr(tpg, "DEVICE", 0)
    check(gpu_pkg)
