"""sarbp -- MI355X-native SAR back-projection (placeholder, filled in below)."""
