"""MI355X-native GKM-WENO flux/update step and Lagrangian particle tracker."""
